// fx_reuse.hpp -- whether a solver set up again may keep its fused-path
// workspace (device vectors sized for the old matrix and replay graphs whose
// kernel arguments point into the old matrix and preconditioner arrays).
//
// Host-only and free of HIP types, so the decision can be unit-checked by a
// plain C++ compiler.  M and P are the matrix and preconditioner handle
// structs; each carries the id `gen` it was given when it was made.
#pragma once

#include <cstdint>

namespace fx {

template <class M, class P>
inline bool solver_reuse_ok(const M* oldA, uint64_t oldAgen, const P* oldPc, uint64_t oldPcGen,
                            const M* A, const P* pc) {
  // Same addresses are not enough: a handle freed and made again usually
  // comes back at the address of the one just freed, with new arrays.
  return oldA != nullptr && oldA == A && oldAgen == A->gen && oldPc != nullptr && oldPc == pc &&
         oldPcGen == pc->gen;
}

}  // namespace fx
