c dropin.h -- the COMMON blocks of the reference's input.h (input.h:19-95),
c declared for a 34 x 18 x 18 grid (NGMX=5, NGMY=NGMZ=4, NGM=5).  The
c block names, member order and types are the interface the library's
c Fortran call forms bind to; the sizes follow input.h's PARAMETER rules.
      implicit none
      INTEGER NGM, MEMLEN, mem
      PARAMETER (NGM=5)
      PARAMETER (MEMLEN=64*((2**(3*NGM)-1)/7) + 64*2**(2*NGM)
     %                  + 192*2**NGM + 72*NGM)
      INTEGER icc(NGM),iu(NGM),ires(NGM),irhs(NGM),ia(NGM),ib(NGM),
     %        ic(NGM),iei(NGM),ng
      DOUBLE PRECISION z(MEMLEN)
      COMMON /pointers/ icc,iu,ires,irhs,ia,ib,ic,iei,ng
      COMMON /memory/ z, mem
      INTEGER NMX, NMY, NMZ
      PARAMETER (NMX=34, NMY=18, NMZ=18)
      DOUBLE PRECISION u(NMX,NMY,NMZ),v(NMX,NMY,NMZ),w(NMX,NMY,NMZ)
      DOUBLE PRECISION S11(NMX,NMY,NMZ),S22(NMX,NMY,NMZ),
     %                 S33(NMX,NMY,NMZ),S12(NMX,NMY,NMZ),
     %                 S13(NMX,NMY,NMZ),S23(NMX,NMY,NMZ)
      COMMON /surfcom/ u,v,w,S11,S22,S33,S12,S13,S23
      DOUBLE PRECISION mut(NMX,NMY,NMZ),lmix(NMX),cmu(NMX,NMY,NMZ),
     & turbst(NMX,NMY,NMZ),ktur(NMX,NMY,NMZ),eps(NMX,NMY,NMZ),
     & u2p(NMX,NMY,NMZ),v2p(NMX,NMY,NMZ),w2p(NMX,NMY,NMZ),
     & uvp(NMX,NMY,NMZ),uwp(NMX,NMY,NMZ),vwp(NMX,NMY,NMZ),
     & hx(NMX,NMY,NMZ),hy(NMX,NMY,NMZ),hz(NMX,NMY,NMZ),
     & ax(NMX,NMY,NMZ),ay(NMX,NMY,NMZ),az(NMX,NMY,NMZ),
     & bx(NMX,NMY,NMZ),by(NMX,NMY,NMZ),bz(NMX,NMY,NMZ)
      COMMON /turbulence/ mut,lmix,turbst,ktur,eps,cmu,
     & u2p,v2p,w2p,uvp,uwp,vwp,hx,hy,hz,ax,ay,az,bx,by,bz
      INTEGER nx,ny,nz
      COMMON /dimension/ nx,ny,nz
      DOUBLE PRECISION ANGLEDIF,MAXCFL,OMEGA,EPSC,EPSURR,tauh
      INTEGER CYCLES,NPRE,NPOST,NTOP,restart,icase,ijet,idrop,
     % tauh2,ibubble,isheet,iturb,nfilter,ibc,tswap,convectype,
     % hswitch,surfacet
      COMMON /DIFF/ ANGLEDIF,EPSC,EPSURR,tauh,tauh2
      COMMON /CYCLING/ MAXCFL,OMEGA,CYCLES,NPRE,NPOST,NTOP,
     % restart,icase,ijet,idrop,ibubble,isheet,iturb,
     % nfilter,ibc,tswap,convectype,hswitch,surfacet
      INTEGER t0,tmax,tstore,ntecplot,tsave,tprint
      COMMON /TIMES/ t0,tmax,tstore,ntecplot,tsave,tprint
      DOUBLE PRECISION tau_SI,timescale,lengthscale,length_SI,
     %     rho1_SI,rho2_SI,visc1_SI,visc2_SI,sigma_SI,grav_SI,
     %     lengthsize,numer,grav1_SI,xvel1_SI,yvel1_SI,zvel1_SI,
     %     xvel2_SI,yvel2_SI,zvel2_SI,radius_si,z0_SI,y0_SI
      COMMON /dinem/ tau_SI,timescale,lengthscale,length_SI,
     %     rho1_SI,rho2_SI,visc1_SI,visc2_SI,sigma_SI,grav_SI,
     %     lengthsize,numer,grav1_SI,xvel1_SI,yvel1_SI,zvel1_SI,
     %     xvel2_SI,yvel2_SI,zvel2_SI,radius_si,z0_SI,y0_SI
      DOUBLE PRECISION tau,time,radius,h,rho1,rho2,mu1,mu2,sigma,g,
     % g1,y0,z0,xvel1,yvel1,zvel1,xvel2,yvel2,zvel2,
     % rhon(NMX,NMY,NMZ),mun(NMX,NMY,NMZ)
      COMMON /PHYSICS/ tau,time,radius,h,rho1,rho2,mu1,mu2,sigma,g,
     % g1,y0,z0,xvel1,yvel1,zvel1,xvel2,yvel2,zvel2,rhon,mun
      DOUBLE PRECISION hxmax(NMX,NMY,NMZ),hxmin(NMX,NMY,NMZ),
     & hymax(NMX,NMY,NMZ),hymin(NMX,NMY,NMZ),
     & hzmax(NMX,NMY,NMZ),hzmin(NMX,NMY,NMZ)
      COMMON /NEWW/ hxmax,hxmin,hymax,hymin,hzmax,hzmin
