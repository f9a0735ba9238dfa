// dg_resident.h -- set-up of the process-wide resident state, shared by the
// C-ABI (dg_api.hip) and the Fortran call forms (dg_fortran.hip).
#pragma once

#include <vector>

#include "dg_ctx.h"

namespace dg {

// The 38 values of inputdata.in in readinput's order (main.f:185-272).
struct Input {
    int nx, ny, nz, tmax, nfilter, cycles;
    double omega;
    int npre, ntop, npost, restart;
    double maxcfl;
    int iturb, convectype, hswitch;
    double xvel1, yvel1, zvel1, xvel2, yvel2, zvel2, rho2, rho1, sigma, visc2, visc1, grav, grav1;
    int icase, ibc;
    double length, radius, y0, z0;
    int tprint, ntecplot, tsave, surfacet;
};

int read_input(const char *path, Input &in);
// nondim (main.f:341-446) into the resident scalars
void nondim(const Input &in, Scalars &s);

// dg_init's work: readinput, memoryaloc, nondim and the droplet IC on the
// device; in_out (optional) receives the parsed input
int init_from_input(const char *path, Input *in_out);
// a fresh resident state of nx x ny x nz with every field zero
int resident_setup(int nx, int ny, int nz);
// CalcH3d's six interface-height arrays (zeroed), once
int alloc_heights(Ctx &c);
// uv2p's six Reynolds-stress arrays (zeroed), once
int alloc_reynolds(Ctx &c);
// recomputes what depends on the physical scalars (after they are set)
void resident_scalars_changed();
bool resident_ready();
// device array of a field id (DG_F_*), nullptr when not ready
double *resident_field(int id);

// The run's inputdata.in values as printout and savedata use them (after a
// restart: with the values iniget read from the file) and its first step.
struct RunInfo {
    bool valid = false;
    Input in{};
    int t0 = 0;
};
RunInfo &run_info();

// iniget (main.f:698-793): header values into `in`, t0, u/v/w/cc records
int read_restart_file(const char *path, Input &in, int *t0, std::vector<double> uvwcc[4]);
// savedata's records (main.f:5414-5467) of host arrays u, v, w, cc
int write_restart_records(const char *path, const Input &in, int t0, const Scalars &s, const double *u,
                          const double *v, const double *w, const double *cc);

// printout's writer (dg_output.hip): waits for queued output and stops its
// thread (dg_finalize); output_active: dg_output_open has opened the files
int output_shutdown();
bool output_active();

// single operators of the time step on resident arrays
int op_swp(Ctx &c, const double *us, double *cc, double *vof1, double *vof2, double *vof3, int d);
int op_convective(Ctx &c, const double *w1, const double *w2, const double *w3);

}  // namespace dg
