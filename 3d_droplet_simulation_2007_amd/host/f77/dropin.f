c dropin.f -- a Fortran-77 program that drives libdroplet_gpu.so through
c the reference's own call forms: implicit interfaces, every argument by
c reference, arrays passed as z(icc(ng)) and the like, state in COMMON.
c The loop is PROGRAM THREE DIM's (main.f:19-73) without printout.
c
c usage: dropin_f77 VARIANT OUT     (run where ./inputdata.in is)
c   VARIANT 1: the loop's calls VOFmethod, CalcH3d (hswitch = 1),
c              kepsmodel, checktime, csf / SGIP / SGIP2 (surfacet 2, 5, 6),
c              solvelpres, as main.f:27-69;
c   VARIANT 2: the same with VOFmethod's body written out (main.f:795-824):
c              the three swp calls in tswap order, bc_c and calcrho, plus
c              the CFL function as printout calls it (main.f:5048).
c Writes tau, time, NPRE, NTOP, NPOST, tswap and cc, u, v, w, p and the
c heights hxmin, hymax, hzmin (raw doubles) to OUT after the last step.
      PROGRAM DROPIN
      INCLUDE 'dropin.h'
      INTEGER t, variant, q, i, j, k
      DOUBLE PRECISION CFL, c1
      EXTERNAL CFL
      CHARACTER*256 arg, outp
c
      call get_command_argument(1, arg)
      read (arg, *) variant
      call get_command_argument(2, outp)
      call dg_init_input()
      do t = t0, tmax
        if (variant .eq. 1) then
          call VOFmethod(t)
        else
          tswap = tswap + 1
          if (tswap .gt. 3) tswap = 1
          if (MOD(tswap,3) .eq. 0) then
            call swp(w,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),3)
            call swp(u,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),1)
            call swp(v,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),2)
          elseif (MOD(tswap,2) .eq. 0) then
            call swp(v,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),2)
            call swp(w,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),3)
            call swp(u,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),1)
          else
            call swp(u,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),1)
            call swp(v,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),2)
            call swp(w,z(icc(ng)),z(iei(ng)),z(ires(ng)),z(irhs(ng)),3)
          endif
          call bc_c(z(icc(ng)))
          call calcrho(z(icc(ng)))
          c1 = CFL()
          write (*,'(a,i6,a,es22.15)') ' step', t, '  CFL', c1
        endif
        if (hswitch.eq.1) call CalcH3d(z(icc(ng)))
        if (iturb.eq.3.or.iturb.eq.4)
     &    call kepsmodel(t,z(iu(ng)),z(ires(ng)))
        call checktime(t)
        if (surfacet.eq.2) then
          call csf(z(icc(ng)),z(iu(ng)),z(ires(ng)),z(irhs(ng)))
        else if (surfacet.eq.5) then
          call SGIP(z(icc(ng)),z(iu(ng)),z(ires(ng)),z(irhs(ng)))
        else if (surfacet.eq.6) then
          call SGIP2(z(icc(ng)),z(iu(ng)),z(ires(ng)),z(irhs(ng)))
        else
          STOP 'capalg'
        endif
        call solvelpres(z(icc(ng)),z(iu(ng)),z(ires(ng)),z(irhs(ng))
     %        ,z(ia(ng)),z(ib(ng)),z(ic(ng)),z(iei(ng)),t)
      enddo
      call dg_sync_to_host(-1)
      open (11, file=outp, access='stream', form='unformatted',
     %      status='replace')
      write (11) tau, time
      write (11) NPRE, NTOP, NPOST, tswap
      write (11) (((z(icc(ng)-1+i+nx*(j-1)+nx*ny*(k-1)),
     %             i=1,nx), j=1,ny), k=1,nz)
      write (11) u, v, w
      write (11) (((z(iu(ng)-1+i+nx*(j-1)+nx*ny*(k-1)),
     %             i=1,nx), j=1,ny), k=1,nz)
      write (11) hxmin, hymax, hzmin
      close (11)
      write (*,'(a,i6,a,es22.15,a,i4)') ' final t=', tmax,
     %   '  tau=', tau, '  NTOP=', NTOP
      call dg_finalize()
      end
