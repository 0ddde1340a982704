// hn_gen.h -- on-device generation of synthetic structures (serving workloads).
#pragma once
#include "hn_engine.h"

namespace hn {

// Fill slot `slot` of `st` (uniform PL size n, 2 contacts on blocks nbl and 1)
// with the k-resolved random slab for seed `seed` at transverse k = (kx, ky);
// with_overlap: S = I + s R (st.smode must be 7), else S = I.
void gen_kslab(DevStruct &st, int slot, unsigned long long seed, double kx, double ky,
               int with_overlap, hipStream_t s);
// overlap scale s of the generated slab for PL size n
double kslab_overlap_scale(int n);

}  // namespace hn
