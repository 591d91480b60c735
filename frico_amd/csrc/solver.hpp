// Factor and solve split into host preparation, device work and host completion.  The *_enqueue
// functions only enqueue device work on the context's stream (no synchronisation, no memset or
// memcpy nodes), so a caller may capture them into a HIP graph and replay it; the C-ABI functions
// frico_factor / frico_solve are compositions of these pieces.
#pragma once
#include "ctx.hpp"

namespace frico {

// Checks the state, allocates the workspace and fills wk.  lookahead=false leaves wk without the
// side stream (required under graph capture).
int factor_prepare(frico_ctx* c, int form, LuWork& wk, bool lookahead);
int factor_enqueue(frico_ctx* c, int form, LuWork& wk);
// Synchronises, reads the panel flags (zero pivot, wait limit) and records the factored state.
// Returns zgetrf's info (> 0: zero pivot) or an error.
int factor_finish(frico_ctx* c, int form, const LuWork& wk);

// Complex64-LU refinement.  refine_prepare allocates; refine_enqueue(first) runs one sweep:
//   first: X = 0, R = B;   then D = LU^{-1} R (complex64), X += D, R = B - Z X,
// and leaves the squared column norms of R in c->norms.  overlap: X += D and R = B - Z X run on
// the context's second stream, 128 rows at a time as the backward substitution finishes them
// (only outside graph recording).
int refine_prepare(frico_ctx* c);
int refine_enqueue(frico_ctx* c, bool first, bool overlap = false);
// Column 2-norms of a device matrix (synchronises).
int col_norms(frico_ctx* c, const c128* X, int64_t cols, std::vector<double>& out);
// Reads c->norms (synchronises) and returns max_j ||R_j|| / ||B_j|| given the B norms.
int read_refine_norms(frico_ctx* c, const std::vector<double>& bn, double* worst);

} // namespace frico
