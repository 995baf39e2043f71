// vtk_io.h -- VTK image I/O without VTK: XML ImageData (.vti, appended
// base64, zlib or raw, UInt32/UInt64 headers) and legacy STRUCTURED_POINTS
// (.vtk, BINARY or ASCII).  Arrays are C order (nx, ny, nz), x slowest; files
// are x fastest (vtkio.cxx:133-143).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ns3d_host {

struct Image {
  int nx = 0, ny = 0, nz = 0, ncomp = 1;
  std::vector<double> data;  // values as double, C order, components last
};

// Reads the first scalar array; throws std::runtime_error on failure.
Image read_image(const std::string& path);

// readvtkfile_uint8 (vtkio.cxx:96-150): voxel == porevalue -> 0 (fluid), else 1.
std::vector<uint8_t> read_binary_image(const std::string& path, int porevalue, int& nx, int& ny, int& nz);

// writeScalarImage / writeVectorImage (vtkio.cxx:161-292): Float64, zlib.
void write_scalar(const std::string& path, const double* a, int nx, int ny, int nz);
void write_vector(const std::string& path, const double* ax, const double* ay, const double* az, int nx,
                  int ny, int nz);
// writeBinaryImage: UInt8.
void write_uint8(const std::string& path, const uint8_t* a, int nx, int ny, int nz);

}  // namespace ns3d_host
