// ini.cpp -- the reference's key = value input grammar (ReadInput,
// navierstokes3D.cxx:266-463) and defaults (InputParams, :162-201).
//
// Grammar: blank lines and lines starting with '#' are skipped; keys are
// case-insensitive; whitespace around key and value is ignored; an inline
// '#' comment starts at a '#' that begins the value or follows whitespace;
// unknown keys warn (with the line number) and are ignored; a malformed
// numeric value or a missing 'input' is fatal.  The reference compares the
// lower-cased key with the mixed-case literal "deltaRMR" (:342), so that key
// can never be set and is reported as unknown; this is kept for drop-in
// behaviour.
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/ns3d_host.h"

namespace {

std::string trim(const std::string& s) {
  const char* ws = " \t\r\n";
  const size_t b = s.find_first_not_of(ws);
  if (b == std::string::npos) return std::string();
  const size_t e = s.find_last_not_of(ws);
  return s.substr(b, e - b + 1);
}

std::string lower(std::string s) {
  for (char& c : s)
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
  return s;
}

std::string upper(std::string s) {
  for (char& c : s)
    if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
  return s;
}

bool parse_double(const std::string& s, double& out) {
  if (s.empty()) return false;
  size_t pos = 0;
  try {
    out = std::stod(s, &pos);
  } catch (...) {
    return false;
  }
  return pos == s.size();
}

bool parse_int(const std::string& s, int& out) {
  if (s.empty()) return false;
  size_t pos = 0;
  try {
    out = std::stoi(s, &pos);
  } catch (...) {
    return false;
  }
  return pos == s.size();
}

std::string strip_comment(const std::string& v) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == '#' && (i == 0 || std::isspace((unsigned char)v[i - 1]))) return trim(v.substr(0, i));
  return trim(v);
}

}  // namespace

extern "C" void ns3d_default_run_params(ns3d_run_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->dx = 1.0; p->dy = -1.0; p->dz = -1.0;
  p->mu = 1e-3; p->rho = 1e3; p->p0 = 1e5; p->gradp = 1e2;
  p->alpha = 0.8; p->beta = 1.0;
  p->vsolvertol = 1e-5; p->psolvertol = 1e-6; p->pcsolvertol = 1e-3; p->simpletol = 1e-4; p->deltaRMR = 1e-4;
  p->maxitsolver = 1000; p->maxitSIMPLE = 500; p->N1 = 2; p->N2 = 2; p->save_step = 20;
  p->porevalue = 0; p->direction = 1; p->border = 0;
  p->method = NS3D_METHOD_SIMPLEC;
  p->verbose = 1;
  p->write_files = 1;
  p->device = -1;
  p->outdir = nullptr;
  p->name = nullptr;
  p->precision = 0;
  p->mixed_min_rows = 0;
  p->inner_tol = 1e-4;
}

extern "C" int ns3d_read_ini(const char* path, ns3d_run_params* p, char* image, int cap) {
  ns3d_default_run_params(p);
  std::ifstream in(path);
  const std::string filename = path;
  if (!in) {
    std::cerr << "error: cannot open input file '" << filename << "'\n";
    return 1;
  }
  bool has_image = false;
  std::string img, method = "SIMPLEC", line;
  std::string pfield = "none", vx = "none", vy = "none", vz = "none";
  size_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) {
      std::cerr << "warning: " << filename << ":" << lineno << " - line ignored (no '=' found)\n";
      continue;
    }
    const std::string key = lower(trim(line.substr(0, eq)));
    const std::string value = strip_comment(line.substr(eq + 1));
    auto bad = [&](const char* what) {
      std::cerr << "error: " << filename << ":" << lineno << " - key '" << key << "' expects a " << what
                << ", got '" << value << "'\n";
      return 1;
    };
    int iv;
    if (key == "input") { img = value; has_image = true; }
    else if (key == "pressurefield") pfield = value;
    else if (key == "vxfield") vx = value;
    else if (key == "vyfield") vy = value;
    else if (key == "vzfield") vz = value;
    else if (key == "dx") { if (!parse_double(value, p->dx)) return bad("number"); }
    else if (key == "dy") { if (!parse_double(value, p->dy)) return bad("number"); }
    else if (key == "dz") { if (!parse_double(value, p->dz)) return bad("number"); }
    else if (key == "mu") { if (!parse_double(value, p->mu)) return bad("number"); }
    else if (key == "rho") { if (!parse_double(value, p->rho)) return bad("number"); }
    else if (key == "p0") { if (!parse_double(value, p->p0)) return bad("number"); }
    else if (key == "gradp") { if (!parse_double(value, p->gradp)) return bad("number"); }
    else if (key == "alpha") { if (!parse_double(value, p->alpha)) return bad("number"); }
    else if (key == "beta") { if (!parse_double(value, p->beta)) return bad("number"); }
    else if (key == "vsolvertol") { if (!parse_double(value, p->vsolvertol)) return bad("number"); }
    else if (key == "psolvertol") { if (!parse_double(value, p->psolvertol)) return bad("number"); }
    else if (key == "pcsolvertol") { if (!parse_double(value, p->pcsolvertol)) return bad("number"); }
    else if (key == "simpletol") { if (!parse_double(value, p->simpletol)) return bad("number"); }
    else if (key == "deltaRMR") { if (!parse_double(value, p->deltaRMR)) return bad("number"); }  // never true
    else if (key == "maxitsolver") { if (!parse_int(value, iv)) return bad("integer"); p->maxitsolver = iv; }
    else if (key == "maxitsimple") { if (!parse_int(value, iv)) return bad("integer"); p->maxitSIMPLE = iv; }
    else if (key == "n1") { if (!parse_int(value, iv)) return bad("integer"); p->N1 = iv; }
    else if (key == "n2") { if (!parse_int(value, iv)) return bad("integer"); p->N2 = iv; }
    else if (key == "save_step") { if (!parse_int(value, iv)) return bad("integer"); p->save_step = iv; }
    else if (key == "porevalue") { if (!parse_int(value, iv)) return bad("integer"); p->porevalue = iv; }
    else if (key == "direction") { if (!parse_int(value, iv)) return bad("integer"); p->direction = iv; }
    else if (key == "border") { if (!parse_int(value, iv)) return bad("integer"); p->border = iv; }
    else if (key == "method") method = value;
    else std::cerr << "warning: " << filename << ":" << lineno << " - unknown key '" << key << "' (ignored)\n";
  }
  if (!has_image) {
    std::cerr << "error: " << filename << " - required key 'input' (image file) is missing\n";
    return 1;
  }
  if (img.empty()) {
    std::cerr << "error: " << filename << " - key 'input' has an empty value\n";
    return 1;
  }
  if (p->dx <= 0.0) { std::cerr << "error: " << filename << " - 'dx' must be > 0\n"; return 1; }
  if (p->mu <= 0.0) { std::cerr << "error: " << filename << " - 'mu' must be > 0\n"; return 1; }
  if (p->rho <= 0.0) { std::cerr << "error: " << filename << " - 'rho' must be > 0\n"; return 1; }
  if (p->maxitsolver <= 0) { std::cerr << "error: " << filename << " - 'maxitsolver' must be > 0\n"; return 1; }
  if (p->maxitSIMPLE <= 0) { std::cerr << "error: " << filename << " - 'maxitSIMPLE' must be > 0\n"; return 1; }
  if (p->save_step <= 0) { std::cerr << "error: " << filename << " - 'save_step' must be > 0\n"; return 1; }
  if (p->porevalue != 0 && p->porevalue != 1) {
    std::cerr << "warning: " << filename << " - 'porevalue' should be 0 or 1, treating '" << p->porevalue
              << "' as 0\n";
    p->porevalue = 0;
  }
  if (p->direction < 1 || p->direction > 3) {
    std::cerr << "error: " << filename << " - 'direction' must be 1 (x), 2 (y) or 3 (z)\n";
    return 1;
  }
  const std::string m = lower(method);
  if (m != "simplec" && m != "simpler" && m != "simplerr" && m != "ideal") {
    std::cerr << "error: " << filename << " - unknown method '" << method
              << "'. Valid values: SIMPLEC, SIMPLER, SIMPLERR, IDEAL\n";
    return 1;
  }
  p->method = m == "simplec" ? NS3D_METHOD_SIMPLEC : m == "simpler" ? NS3D_METHOD_SIMPLER
            : m == "simplerr" ? NS3D_METHOD_SIMPLERR : NS3D_METHOD_IDEAL;
  if (p->dy <= 0.0) p->dy = p->dx;
  if (p->dz <= 0.0) p->dz = p->dx;
  // the image path and the warm-start field paths travel in one buffer:
  // "image\0pfield\0vx\0vy\0vz\0"
  const std::string packed = img + '\0' + pfield + '\0' + vx + '\0' + vy + '\0' + vz + '\0';
  if ((int)packed.size() > cap) {
    std::cerr << "error: file names too long\n";
    return 1;
  }
  std::memcpy(image, packed.data(), packed.size());
  printf("=== Input parameters ===\n");
  printf("method         : %s\n", upper(method).c_str());
  printf("image          : %s\n", img.c_str());
  printf("dx, dy, dz (m) : %g, %g, %g\n", p->dx, p->dy, p->dz);
  printf("mu (Pa.s)      : %g\n", p->mu);
  printf("rho (kg/m^3)   : %g\n", p->rho);
  printf("p0 (Pa)        : %g\n", p->p0);
  printf("gradp (Pa/m)   : %g\n", p->gradp);
  printf("alpha, beta    : %g, %g\n", p->alpha, p->beta);
  printf("simpletol      : %g\n", p->simpletol);
  printf("maxitSIMPLE    : %d\n", p->maxitSIMPLE);
  printf("maxitsolver    : %d\n", p->maxitsolver);
  printf("save_step      : %d\n", p->save_step);
  printf("porevalue      : %d\n", p->porevalue);
  printf("direction      : %d\n", p->direction);
  printf("border         : %d\n", p->border);
  printf("======================\n");
  return 0;
}
