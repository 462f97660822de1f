// ASCII VTI output (io/vtiWriter.cuh:60-145): ImageData, five Float32 point arrays rhor rhob ux
// uy uz, x fastest, one text row per (z, y) with every value printed "%.8e " style, then the
// empty CellData block. Byte-identical to the reference writer for the same arrays.
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <string>
#include <system_error>
#include <vector>

#include "cglbm.h"
#include "state.hpp"
#include "vti.hpp"

namespace cglbm {

int write_vti_arrays(const char *path, int nx, int ny, int nz, const float *const fields[5]) {
    if (!path || !*path) return set_error(CGLBM_EINVAL, "null or empty VTI path");
    if (nx <= 0 || ny <= 0 || nz <= 0) return set_error(CGLBM_EINVAL, "VTI extent must be positive");
    for (int f = 0; f < 5; ++f)
        if (!fields[f]) return set_error(CGLBM_EINVAL, "null VTI field pointer");
    const std::filesystem::path file(path);
    if (file.has_parent_path()) {  // write_vti creates missing folders (io/vtiWriter.cuh:96-99)
        std::error_code ec;
        std::filesystem::create_directories(file.parent_path(), ec);
        if (ec) return set_error(CGLBM_EIO, ("cannot create the folder of " + file.string()).c_str());
    }
    std::ofstream out(file);
    if (!out) return set_error(CGLBM_EIO, ("cannot open VTI file for writing: " + file.string()).c_str());
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1
        << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    out << "    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n";
    out << "      <PointData Scalars=\"rhor\">\n";
    static const char *const names[5] = {"rhor", "rhob", "ux", "uy", "uz"};
    const int64_t row = nx, plane = static_cast<int64_t>(nx) * ny;
    for (int f = 0; f < 5; ++f) {
        out << "        <DataArray type=\"Float32\" Name=\"" << names[f] << "\" format=\"ascii\">\n";
        out << std::scientific << std::setprecision(8);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                const float *p = fields[f] + plane * z + row * y;
                for (int x = 0; x < nx; ++x) out << p[x] << " ";
                out << "\n";
            }
        out << "        </DataArray>\n";
    }
    out << "      </PointData>\n";
    out << "      <CellData>\n";
    out << "      </CellData>\n";
    out << "    </Piece>\n";
    out << "  </ImageData>\n";
    out << "</VTKFile>\n";
    out.flush();
    if (!out) return set_error(CGLBM_EIO, ("error while finalising VTI file: " + file.string()).c_str());
    return CGLBM_OK;
}

}  // namespace cglbm

extern "C" int cglbm_write_vti_host(const char *path, int nx, int ny, int nz, const float *rhor,
                                    const float *rhob, const float *ux, const float *uy, const float *uz) {
    const float *const f[5] = {rhor, rhob, ux, uy, uz};
    return cglbm::write_vti_arrays(path, nx, ny, nz, f);
}
