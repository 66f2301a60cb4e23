// lbfgs_api.h — the l_bfgs driver as seen by the engine (no device code).
#pragma once
#include <functional>

#include "common.h"

struct LbfgsDev;

// Misfit/gradient evaluation at the FP32 model x (the reference's fctgdt).
typedef std::function<int(const float *x, float *g, double *f)> LbfgsFct;

// Extension points of the driver (the learned preconditioner uses them).
struct LbfgsHooks {
    virtual ~LbfgsHooks() {}
    virtual void attach(LbfgsDev &d) {}
    virtual void before_step(LbfgsDev &d, double fscaled) {}
    virtual void accepted(LbfgsDev &d, int iteration) {}
    // after each LBFGS call; may replace the IFLAG (e.g. restart a rejected search)
    virtual int after_step(LbfgsDev &d, int iflag) { return iflag; }
    // the evaluation rejected the trial model of a search along the hook's own
    // direction; returns the IFLAG of the search that replaces it (< 0: none)
    virtual int rejected_trial(LbfgsDev &d) { return -1; }
    virtual void report(sfd_lbfgs_report *rep) {}
    // compute_fct_gdt's illumination at the newest evaluated model (host, n values)
    virtual bool wants_illumination() const { return false; }
    virtual void set_illumination(const float *illum, long n) {}
};

// Inversion.f90:29-189 l_bfgs; returns the final IFLAG.
int lbfgs_drive(const sfd_lbfgs_opts &o, long n, float *x, const float *mask, const double *vpinit,
                const LbfgsFct &fct, sfd_lbfgs_report *rep, LbfgsHooks *hooks);
