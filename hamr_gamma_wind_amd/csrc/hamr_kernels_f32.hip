// hamr_kernels_f32.hip -- the FP32 reconstruction, HLL flux and EMF kernels of
// the mixed step, in a unit of their own (built concurrently with the FP64 one).
#include "hamr_recon.h"

namespace hamr {

void launch_flux_f32(const KParams &P, int dir, const Launch &L, const double *p, const double *ps, float *F,
                     unsigned long long *dtmin, hipStream_t s)
{
    recon_flux<float>(P, dir, L, p, ps, F, dtmin, s);
}

void launch_emf_centre_f32(const KParams &P, const Launch &L, const double *p, float *Ec, hipStream_t s)
{
    recon_emf_centre<float>(P, L, p, Ec, s);
}

void launch_emf_corner_f32(const KParams &P, const Launch &L, const float *Ec, const float *F1, const float *F2,
                           const float *F3, float *E, hipStream_t s)
{
    recon_emf_corner<float>(P, L, Ec, F1, F2, F3, E, s);
}

}  // namespace hamr
