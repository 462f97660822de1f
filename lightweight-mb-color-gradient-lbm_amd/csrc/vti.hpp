// VTI text writer shared by cglbm_write_vti (domain state) and cglbm_write_vti_host (host arrays).
#pragma once

namespace cglbm {
// fields: rhor, rhob, ux, uy, uz, each nx*ny*nz floats, x fastest.
int write_vti_arrays(const char *path, int nx, int ny, int nz, const float *const fields[5]);
}  // namespace cglbm
