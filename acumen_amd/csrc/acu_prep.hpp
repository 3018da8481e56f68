// acu_prep.hpp -- projection -> sinogram preprocessing (prep.hip), the
// restatement of tomography/projections_to_sinogram_hpc.F90:231-487.
#pragma once

#include "acu_internal.hpp"

namespace acu {

// Geometry of the resampling of one (theta, shift) trial or correction:
// Nw = n + add_nx, mx = (Nw-1)/2, my = (n-1)/2, S = n, c / s = cos / sin of
// theta in FP32 (the reference's expressions).
struct PrepGeom {
  int n, add_nx;
  float mx, my, S, shift, c, s;
};

PrepGeom prep_geom(int n, int add_nx, float theta, float shift);
void launch_median27(int n, const float* l, const float* c, const float* r, float* out, hipStream_t s);
void launch_correct(const PrepGeom& g, const float* imgs, int count, int job0, int nfiles, float* sino,
                    hipStream_t s);
int mirror_blocks(const PrepGeom& g);
void launch_mirror(const PrepGeom& g, const float* p0, const float* p180, double* part, double* out,
                   hipStream_t s);
// FITS ingest normalisation (fits_to_hpc.F90:196-248)
void launch_mean(const float* imgs, int count, size_t npix, float* out, hipStream_t s);
void launch_normalize(const float* imgs, int count, size_t npix, const float* ob, float* out, hipStream_t s);

}  // namespace acu
