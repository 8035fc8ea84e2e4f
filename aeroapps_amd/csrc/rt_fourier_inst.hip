// One (streams, Stokes) instantiation of the Fourier-term kernels.  The build
// compiles this file once per supported pair with AA_PART_N / AA_PART_S set,
// so the instantiations compile in parallel; rt_fourier_N_S is the entry the
// dispatcher calls.
#include "rt_fourier.h"

#if !defined(AA_PART_N) || !defined(AA_PART_S)
#error "compile with -DAA_PART_N=<streams> -DAA_PART_S=<1|3>"
#endif

#define AA_CAT2(a, b, c, d) a##b##c##d
#define AA_CAT(a, b, c, d) AA_CAT2(a, b, c, d)

namespace aa {
int AA_CAT(rt_fourier_, AA_PART_N, _, AA_PART_S)(const RTWork& W, int nitems, int* const lists[2], int* counts,
                                                  hipStream_t st) {
  return fourier_loop<AA_PART_N, AA_PART_S>(W, nitems, lists, counts, st);
}
}  // namespace aa
