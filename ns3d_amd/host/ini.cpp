// ini.cpp -- the .ini input of the NS3D command line: key = value lines,
// the keys and defaults of the reference (InputParams and ReadInput,
// navierstokes3D.cxx:162-201, :266-463), and the messages it prints.
//
// Design: every key is one row of a table that says how its value is read
// (real, integer or text) and where it is stored; a line is split by a small
// scanner over [begin, end) character ranges; the checks after the last line
// are a list of rules, each with its message.  What a user sees is the
// reference's behaviour:
//  * blank lines and lines whose first non-blank character is '#' are
//    skipped, and so are lines without '=' (with a warning);
//  * the key (left of the first '=') is matched in lower case; the value is
//    the right-hand side without surrounding blanks and without an inline
//    comment, which starts at a '#' that opens the value or follows a
//    white-space character;
//  * a value that is not entirely a number (or an integer) is fatal;
//    unknown keys warn and are ignored; 'input' is required;
//  * the deltaRMR row keeps its mixed-case spelling, so a lower-cased key
//    never matches it: the key is reported as unknown, as in the reference
//    (navierstokes3D.cxx:342).
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/ns3d_host.h"

namespace {

// ---- character ranges --------------------------------------------------------
struct Span {
  const char* b;
  const char* e;
  bool empty() const { return b == e; }
  std::string str() const { return std::string(b, e); }
};

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

Span strip(Span s) {
  while (s.b < s.e && blank(*s.b)) ++s.b;
  while (s.e > s.b && blank(s.e[-1])) --s.e;
  return s;
}

// The value part of a line: blanks stripped, then cut at the first '#' that
// opens it or follows white space (a '#' inside a file name stays).
Span value_of(Span rhs) {
  Span v = strip(rhs);
  for (const char* c = v.b; c < v.e; ++c)
    if (*c == '#' && (c == v.b || std::isspace((unsigned char)c[-1]))) return strip(Span{v.b, c});
  return v;
}

std::string ascii_case(std::string s, bool up) {
  for (char& c : s) {
    if (!up && c >= 'A' && c <= 'Z') c = (char)(c + ('a' - 'A'));
    if (up && c >= 'a' && c <= 'z') c = (char)(c - ('a' - 'A'));
  }
  return s;
}

// Whole-string numbers.  Overflow and underflow (ERANGE) are errors, as for
// std::stod / std::stoi.
bool read_real(const std::string& s, double* out) {
  if (s.empty()) return false;
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || errno == ERANGE || end != s.c_str() + s.size()) return false;
  *out = v;
  return true;
}

bool read_integer(const std::string& s, int32_t* out) {
  if (s.empty()) return false;
  errno = 0;
  char* end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || errno == ERANGE || v < INT_MIN || v > INT_MAX || end != s.c_str() + s.size())
    return false;
  *out = (int32_t)v;
  return true;
}

// ---- the key table -------------------------------------------------------------
enum class Kind { Real, Integer, Text };

struct Row {
  const char* spelling;  // compared with the lower-cased key
  Kind kind;
  double* real;
  int32_t* integer;
  std::string* text;
};

// What the parse collects besides ns3d_run_params.
struct Texts {
  std::string image, method = "SIMPLEC";
  std::string pfield = "none", vx = "none", vy = "none", vz = "none";
  bool has_image = false;
};

std::vector<Row> key_table(ns3d_run_params* p, Texts* t) {
  auto R = [](const char* k, double* d) { return Row{k, Kind::Real, d, nullptr, nullptr}; };
  auto I = [](const char* k, int32_t* i) { return Row{k, Kind::Integer, nullptr, i, nullptr}; };
  auto S = [](const char* k, std::string* s) { return Row{k, Kind::Text, nullptr, nullptr, s}; };
  return {
      S("input", &t->image),          S("pressurefield", &t->pfield), S("vxfield", &t->vx),
      S("vyfield", &t->vy),           S("vzfield", &t->vz),           R("dx", &p->dx),
      R("dy", &p->dy),                R("dz", &p->dz),                R("mu", &p->mu),
      R("rho", &p->rho),              R("p0", &p->p0),                R("gradp", &p->gradp),
      R("alpha", &p->alpha),          R("beta", &p->beta),            R("vsolvertol", &p->vsolvertol),
      R("psolvertol", &p->psolvertol), R("pcsolvertol", &p->pcsolvertol), R("simpletol", &p->simpletol),
      R("deltaRMR", &p->deltaRMR),    I("maxitsolver", &p->maxitsolver), I("maxitsimple", &p->maxitSIMPLE),
      I("n1", &p->N1),                I("n2", &p->N2),                I("save_step", &p->save_step),
      I("porevalue", &p->porevalue),  I("direction", &p->direction),  I("border", &p->border),
      S("method", &t->method),
  };
}

const char* kMethods[4] = {"simplec", "simpler", "simplerr", "ideal"};

int method_id(const std::string& m) {
  const std::string l = ascii_case(m, false);
  for (int q = 0; q < 4; ++q)
    if (l == kMethods[q]) return q;  // NS3D_METHOD_* order
  return -1;
}

// One fatal rule after parsing: the check and the text after "error: <file> - ".
struct Rule {
  bool (*broken)(const ns3d_run_params&, const Texts&);
  const char* message;
};

const Rule kRules[] = {
    {[](const ns3d_run_params&, const Texts& t) { return !t.has_image; },
     "required key 'input' (image file) is missing"},
    {[](const ns3d_run_params&, const Texts& t) { return t.image.empty(); }, "key 'input' has an empty value"},
    {[](const ns3d_run_params& p, const Texts&) { return p.dx <= 0.0; }, "'dx' must be > 0"},
    {[](const ns3d_run_params& p, const Texts&) { return p.mu <= 0.0; }, "'mu' must be > 0"},
    {[](const ns3d_run_params& p, const Texts&) { return p.rho <= 0.0; }, "'rho' must be > 0"},
    {[](const ns3d_run_params& p, const Texts&) { return p.maxitsolver <= 0; }, "'maxitsolver' must be > 0"},
    {[](const ns3d_run_params& p, const Texts&) { return p.maxitSIMPLE <= 0; }, "'maxitSIMPLE' must be > 0"},
    {[](const ns3d_run_params& p, const Texts&) { return p.save_step <= 0; }, "'save_step' must be > 0"},
};

void print_summary(const ns3d_run_params& p, const Texts& t) {
  std::printf("=== Input parameters ===\n");
  std::printf("method         : %s\n", ascii_case(t.method, true).c_str());
  std::printf("image          : %s\n", t.image.c_str());
  std::printf("dx, dy, dz (m) : %g, %g, %g\n", p.dx, p.dy, p.dz);
  const struct { const char* label; double v; } reals[] = {
      {"mu (Pa.s)      ", p.mu}, {"rho (kg/m^3)   ", p.rho}, {"p0 (Pa)        ", p.p0},
      {"gradp (Pa/m)   ", p.gradp}};
  for (const auto& r : reals) std::printf("%s: %g\n", r.label, r.v);
  std::printf("alpha, beta    : %g, %g\n", p.alpha, p.beta);
  std::printf("simpletol      : %g\n", p.simpletol);
  const struct { const char* label; int v; } ints[] = {
      {"maxitSIMPLE    ", p.maxitSIMPLE}, {"maxitsolver    ", p.maxitsolver}, {"save_step      ", p.save_step},
      {"porevalue      ", p.porevalue},   {"direction      ", p.direction},   {"border         ", p.border}};
  for (const auto& r : ints) std::printf("%s: %d\n", r.label, r.v);
  std::printf("======================\n");
  std::fflush(stdout);
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
  p->inner_tol = 1e-4;
  p->mg_min_rows = 200000;
}

extern "C" int ns3d_read_ini(const char* path, ns3d_run_params* p, char* image, int cap) {
  ns3d_default_run_params(p);
  const std::string file = path;
  std::ifstream in(file);
  if (!in) {
    std::cerr << "error: cannot open input file '" << file << "'\n";
    return 1;
  }
  Texts t;
  const std::vector<Row> table = key_table(p, &t);
  std::string text;
  for (size_t lineno = 1; std::getline(in, text); ++lineno) {
    const Span line = strip(Span{text.data(), text.data() + text.size()});
    if (line.empty() || *line.b == '#') continue;
    const char* eq = static_cast<const char*>(std::memchr(line.b, '=', (size_t)(line.e - line.b)));
    if (!eq) {
      std::cerr << "warning: " << file << ":" << lineno << " - line ignored (no '=' found)\n";
      continue;
    }
    const std::string key = ascii_case(strip(Span{line.b, eq}).str(), false);
    const std::string value = value_of(Span{eq + 1, line.e}).str();
    const Row* row = nullptr;
    for (const Row& r : table)
      if (key == r.spelling) { row = &r; break; }
    if (!row) {
      std::cerr << "warning: " << file << ":" << lineno << " - unknown key '" << key << "' (ignored)\n";
      continue;
    }
    bool ok = true;
    switch (row->kind) {
      case Kind::Real: ok = read_real(value, row->real); break;
      case Kind::Integer: ok = read_integer(value, row->integer); break;
      case Kind::Text: *row->text = value; break;
    }
    if (row->text == &t.image) t.has_image = true;
    if (!ok) {
      std::cerr << "error: " << file << ":" << lineno << " - key '" << key << "' expects a "
                << (row->kind == Kind::Real ? "number" : "integer") << ", got '" << value << "'\n";
      return 1;
    }
  }
  for (const Rule& r : kRules)
    if (r.broken(*p, t)) {
      std::cerr << "error: " << file << " - " << r.message << "\n";
      return 1;
    }
  if (p->porevalue != 0 && p->porevalue != 1) {
    std::cerr << "warning: " << file << " - 'porevalue' should be 0 or 1, treating '" << p->porevalue
              << "' as 0\n";
    p->porevalue = 0;
  }
  if (p->direction < 1 || p->direction > 3) {
    std::cerr << "error: " << file << " - 'direction' must be 1 (x), 2 (y) or 3 (z)\n";
    return 1;
  }
  const int m = method_id(t.method);
  if (m < 0) {
    std::cerr << "error: " << file << " - unknown method '" << t.method
              << "'. Valid values: SIMPLEC, SIMPLER, SIMPLERR, IDEAL\n";
    return 1;
  }
  p->method = m;
  if (p->dy <= 0.0) p->dy = p->dx;
  if (p->dz <= 0.0) p->dz = p->dx;
  // the image path and the warm-start field paths travel in one buffer:
  // "image\0pfield\0vx\0vy\0vz\0"
  std::string packed;
  for (const std::string* s : {&t.image, &t.pfield, &t.vx, &t.vy, &t.vz}) {
    packed += *s;
    packed += '\0';
  }
  if ((int)packed.size() > cap) {
    std::cerr << "error: file names too long\n";
    return 1;
  }
  std::memcpy(image, packed.data(), packed.size());
  print_summary(*p, t);
  return 0;
}
