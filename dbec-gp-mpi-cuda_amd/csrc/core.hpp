// core.hpp - entry points of the step and the observables for other
// translation units of the library (the request loop of serve.hip); not part
// of the public ABI.
#pragma once

#include <string>

#include "internal.hpp"

namespace dbec {

// The single solve held by c (its own buffers, trap and couplings).
Solves core_solves(const dbec_ctx *c, bool use_scale);
// One split step of the solves v; c supplies grid, tables and stream.  In
// imaginary time the step leaves the norm in the column partials: set v.fold
// on the next step (2 when this step ran with fuse_next, else 1), or call
// core_norm.  z_done: the previous step ran with fuse_next (its z sweep did
// this step's forward z pass); fuse_next only where core_fuse_z(c) and
// another step follows.
int core_step(dbec_ctx *c, const Solves &v, bool z_done, bool fuse_next);
bool core_fuse_z(const dbec_ctx *c);
int core_norm(dbec_ctx *c, const Solves &v);
// Observables of c's psi, asynchronous on c->stream, results in c->b.scal.
// Imaginary time first applies the pending 1/norm (scal[SC_SCALE] -> psi,
// then scal[SC_SCALE] = 1).  what & OBS_MUEN: mu, E; what & OBS_RMS: norm and
// rms moments.
enum { OBS_MUEN = 1, OBS_RMS = 2 };
int core_observe_async(dbec_ctx *c, int what);
// Sets c's error text and returns code.
int core_fail(dbec_ctx *c, int code, const std::string &msg);
size_t core_psi_elem(const dbec_ctx *c);
size_t core_real_elem(const dbec_ctx *c);

}  // namespace dbec
