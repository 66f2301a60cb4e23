// cnn_api.h — the learned preconditioner as seen by the engine.
#pragma once
#include "lbfgs_api.h"

LbfgsHooks *make_cnn_precond(int nz, int nx, const sfd_precond_opts &o);
