// vtk_io.cpp -- see vtk_io.h.
#include "vtk_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>

#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace ns3d_host {

namespace {

const char* kB64 = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

std::string b64_encode(const uint8_t* p, size_t n) {
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
    out += kB64[v >> 18]; out += kB64[(v >> 12) & 63]; out += kB64[(v >> 6) & 63]; out += kB64[v & 63];
  }
  if (i + 1 == n) {
    uint32_t v = (uint32_t)p[i] << 16;
    out += kB64[v >> 18]; out += kB64[(v >> 12) & 63]; out += "==";
  } else if (i + 2 == n) {
    uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8;
    out += kB64[v >> 18]; out += kB64[(v >> 12) & 63]; out += kB64[(v >> 6) & 63]; out += '=';
  }
  return out;
}

std::vector<uint8_t> b64_decode(const char* s, size_t n) {
  int8_t lut[256];
  std::memset(lut, -1, sizeof(lut));
  for (int q = 0; q < 64; ++q) lut[(uint8_t)kB64[q]] = (int8_t)q;
  std::vector<uint8_t> out;
  out.reserve(n / 4 * 3);
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = 0; i < n; ++i) {
    const int8_t v = lut[(uint8_t)s[i]];
    if (v < 0) continue;  // padding and whitespace
    acc = acc << 6 | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((uint8_t)(acc >> bits));
    }
  }
  return out;
}

std::string attr(const std::string& tag, const std::string& name) {
  const std::string key = name + "=\"";
  size_t p = tag.find(key);
  if (p == std::string::npos) return "";
  p += key.size();
  return tag.substr(p, tag.find('"', p) - p);
}

size_t type_size(const std::string& t) {
  if (t == "UInt8" || t == "Int8") return 1;
  if (t == "UInt16" || t == "Int16") return 2;
  if (t == "UInt32" || t == "Int32" || t == "Float32") return 4;
  if (t == "UInt64" || t == "Int64" || t == "Float64") return 8;
  throw std::runtime_error("unsupported VTK data type " + t);
}

double load_as_double(const uint8_t* p, const std::string& t) {
  if (t == "UInt8") return *p;
  if (t == "Int8") return (int8_t)*p;
  if (t == "UInt16") { uint16_t v; std::memcpy(&v, p, 2); return v; }
  if (t == "Int16") { int16_t v; std::memcpy(&v, p, 2); return v; }
  if (t == "UInt32") { uint32_t v; std::memcpy(&v, p, 4); return v; }
  if (t == "Int32") { int32_t v; std::memcpy(&v, p, 4); return v; }
  if (t == "Float32") { float v; std::memcpy(&v, p, 4); return v; }
  if (t == "UInt64") { uint64_t v; std::memcpy(&v, p, 8); return (double)v; }
  if (t == "Int64") { int64_t v; std::memcpy(&v, p, 8); return (double)v; }
  double v;
  std::memcpy(&v, p, 8);
  return v;
}

// File order (x fastest) -> C order (x slowest).
void to_c_order(const std::vector<uint8_t>& raw, const std::string& type, Image& im) {
  const size_t ts = type_size(type);
  const size_t n = (size_t)im.nx * im.ny * im.nz;
  if (raw.size() < n * im.ncomp * ts) throw std::runtime_error("VTK payload shorter than the extent");
  im.data.assign(n * im.ncomp, 0.0);
  for (int z = 0; z < im.nz; ++z)
    for (int y = 0; y < im.ny; ++y)
      for (int x = 0; x < im.nx; ++x) {
        const size_t f = ((size_t)z * im.ny + y) * im.nx + x;
        const size_t c = ((size_t)x * im.ny + y) * im.nz + z;
        for (int q = 0; q < im.ncomp; ++q)
          im.data[c * im.ncomp + q] = load_as_double(&raw[(f * im.ncomp + q) * ts], type);
      }
}

Image read_vti(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::string s((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const size_t app = s.find("<AppendedData");
  if (app == std::string::npos) throw std::runtime_error(path + ": only appended .vti data is supported");
  const std::string head = s.substr(0, app);
  Image im;
  {
    std::istringstream ext(attr(head, "WholeExtent"));
    int e[6];
    for (int& v : e) ext >> v;
    im.nx = e[1] - e[0] + 1; im.ny = e[3] - e[2] + 1; im.nz = e[5] - e[4] + 1;
  }
  const size_t da = head.find("<DataArray");
  const std::string tag = head.substr(da, head.find('>', da) - da);
  const std::string type = attr(tag, "type");
  const std::string nc = attr(tag, "NumberOfComponents");
  im.ncomp = nc.empty() ? 1 : std::stoi(nc);
  const bool compressed = head.find("vtkZLibDataCompressor") != std::string::npos;
  const bool h64 = attr(head, "header_type") == "UInt64";
  const size_t hs = h64 ? 8 : 4;
  const size_t u = s.find('_', s.find('>', app)) + 1;
  const size_t e = s.find("</AppendedData>", u);
  std::string b64;
  b64.reserve(e - u);
  for (size_t i = u; i < e; ++i)
    if (!isspace((unsigned char)s[i])) b64 += s[i];
  auto hval = [&](const std::vector<uint8_t>& h, size_t q) {
    uint64_t v = 0;
    std::memcpy(&v, &h[q * hs], hs);
    return v;
  };
  std::vector<uint8_t> payload;
  if (compressed) {
    const std::vector<uint8_t> first = b64_decode(b64.data(), std::min<size_t>(b64.size(), 12));
    const uint64_t nb = hval(first, 0);
    const size_t hbytes = (3 + nb) * hs, hchars = 4 * ((hbytes + 2) / 3);
    const std::vector<uint8_t> h = b64_decode(b64.data(), hchars);
    const uint64_t bsize = hval(h, 1), last = hval(h, 2);
    const std::vector<uint8_t> data = b64_decode(b64.data() + hchars, b64.size() - hchars);
    size_t off = 0;
    for (uint64_t b = 0; b < nb; ++b) {
      const uint64_t cs = hval(h, 3 + b);
      const uint64_t us = (b + 1 == nb && last) ? last : bsize;
      std::vector<uint8_t> blk(us);
      uLongf dl = (uLongf)us;
      if (uncompress(blk.data(), &dl, &data[off], (uLong)cs) != Z_OK) throw std::runtime_error(path + ": zlib error");
      payload.insert(payload.end(), blk.begin(), blk.begin() + dl);
      off += cs;
    }
  } else {
    const std::vector<uint8_t> all = b64_decode(b64.data(), b64.size());
    uint64_t n = 0;
    std::memcpy(&n, all.data(), hs);
    payload.assign(all.begin() + hs, all.begin() + hs + n);
  }
  to_c_order(payload, type, im);
  return im;
}

Image read_legacy_vtk(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::string line;
  std::getline(f, line);  // # vtk DataFile
  std::getline(f, line);  // title
  std::string mode;
  std::getline(f, mode);
  const bool binary = mode.find("BINARY") != std::string::npos;
  Image im;
  std::string type;
  while (f >> line) {
    if (line == "DIMENSIONS") f >> im.nx >> im.ny >> im.nz;
    else if (line == "SCALARS") {
      std::string name;
      f >> name >> type;
      std::getline(f, line);
      if (!line.empty() && line.find_first_not_of(" \t\r") != std::string::npos) im.ncomp = std::stoi(line);
    } else if (line == "LOOKUP_TABLE") {
      std::getline(f, line);
      break;
    }
  }
  std::string t = type == "unsigned_char" ? "UInt8" : type == "char" ? "Int8" : type == "short" ? "Int16"
                : type == "unsigned_short" ? "UInt16" : type == "int" ? "Int32" : type == "float" ? "Float32"
                : "Float64";
  const size_t n = (size_t)im.nx * im.ny * im.nz * im.ncomp;
  std::vector<uint8_t> raw(n * type_size(t));
  if (binary) {
    f.read((char*)raw.data(), (std::streamsize)raw.size());
    // legacy binary files are big-endian
    const size_t ts = type_size(t);
    for (size_t i = 0; i < n && ts > 1; ++i) std::reverse(&raw[i * ts], &raw[i * ts] + ts);
  } else {
    for (size_t i = 0; i < n; ++i) {
      double v;
      f >> v;
      if (t == "UInt8") raw[i] = (uint8_t)v;
      else throw std::runtime_error(path + ": ASCII legacy VTK supported for unsigned_char only");
    }
  }
  to_c_order(raw, t, im);
  return im;
}

void write_vti(const std::string& path, const std::string& type, size_t ts, int ncomp,
               const std::vector<uint8_t>& payload, int nx, int ny, int nz, double lo, double hi) {
  const size_t block = 32768;
  std::vector<std::vector<uint8_t>> comp;
  std::vector<uint32_t> header;
  const size_t nb = payload.empty() ? 1 : (payload.size() + block - 1) / block;
  header.push_back((uint32_t)nb);
  header.push_back((uint32_t)block);
  const size_t lastsz = payload.size() - (nb - 1) * block;
  header.push_back((uint32_t)(lastsz == block ? 0 : lastsz));
  for (size_t b = 0; b < nb; ++b) {
    const size_t off = b * block, len = std::min(block, payload.size() - off);
    uLongf cl = compressBound((uLong)len);
    std::vector<uint8_t> c(cl);
    if (compress2(c.data(), &cl, payload.data() + off, (uLong)len, Z_DEFAULT_COMPRESSION) != Z_OK)
      throw std::runtime_error("zlib compression failed");
    c.resize(cl);
    header.push_back((uint32_t)cl);
    comp.push_back(std::move(c));
  }
  std::vector<uint8_t> body;
  for (auto& c : comp) body.insert(body.end(), c.begin(), c.end());
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  (void)ts;
  f << "<?xml version=\"1.0\"?>\n"
    << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\" "
       "compressor=\"vtkZLibDataCompressor\">\n"
    << "  <ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1
    << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\" Direction=\"1 0 0 0 1 0 0 0 1\">\n"
    << "    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n"
    << "      <PointData Scalars=\"ImageScalars\">\n"
    << "        <DataArray type=\"" << type << "\" Name=\"ImageScalars\""
    << (ncomp > 1 ? " NumberOfComponents=\"" + std::to_string(ncomp) + "\"" : std::string())
    << " format=\"appended\" RangeMin=\"" << lo << "\" RangeMax=\"" << hi << "\" offset=\"0\"/>\n"
    << "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n"
    << "  <AppendedData encoding=\"base64\">\n   _"
    << b64_encode((const uint8_t*)header.data(), header.size() * 4) << b64_encode(body.data(), body.size())
    << "\n  </AppendedData>\n</VTKFile>\n";
}

}  // namespace

Image read_image(const std::string& path) {
  if (path.size() >= 3 && path.compare(path.size() - 3, 3, "vti") == 0) return read_vti(path);
  if (path.size() >= 3 && path.compare(path.size() - 3, 3, "vtk") == 0) return read_legacy_vtk(path);
  throw std::runtime_error("unsupported file extension (expected .vti or .vtk): " + path);
}

std::vector<uint8_t> read_binary_image(const std::string& path, int porevalue, int& nx, int& ny, int& nz) {
  Image im = read_image(path);
  nx = im.nx; ny = im.ny; nz = im.nz;
  std::vector<uint8_t> out((size_t)nx * ny * nz);
  for (size_t i = 0; i < out.size(); ++i)
    out[i] = ((uint8_t)im.data[i * im.ncomp] == (uint8_t)porevalue) ? 0 : 1;
  return out;
}

void write_scalar(const std::string& path, const double* a, int nx, int ny, int nz) {
  const size_t n = (size_t)nx * ny * nz;
  std::vector<uint8_t> payload(n * 8);
  double lo = n ? a[0] : 0, hi = lo;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const double v = a[((size_t)x * ny + y) * nz + z];
        lo = std::min(lo, v); hi = std::max(hi, v);
        std::memcpy(&payload[(((size_t)z * ny + y) * nx + x) * 8], &v, 8);
      }
  write_vti(path, "Float64", 8, 1, payload, nx, ny, nz, lo, hi);
}

void write_vector(const std::string& path, const double* ax, const double* ay, const double* az, int nx, int ny,
                  int nz) {
  const size_t n = (size_t)nx * ny * nz;
  std::vector<uint8_t> payload(n * 24);
  double lo = 0, hi = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t c = ((size_t)x * ny + y) * nz + z, f = ((size_t)z * ny + y) * nx + x;
        const double v[3] = {ax[c], ay[c], az[c]};
        const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        lo = (f == 0) ? m : std::min(lo, m);
        hi = std::max(hi, m);
        std::memcpy(&payload[f * 24], v, 24);
      }
  write_vti(path, "Float64", 8, 3, payload, nx, ny, nz, lo, hi);
}

void write_uint8(const std::string& path, const uint8_t* a, int nx, int ny, int nz) {
  std::vector<uint8_t> payload((size_t)nx * ny * nz);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) payload[((size_t)z * ny + y) * nx + x] = a[((size_t)x * ny + y) * nz + z];
  write_vti(path, "UInt8", 1, 1, payload, nx, ny, nz, 0, 1);
}

}  // namespace ns3d_host
