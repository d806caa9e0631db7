// S-parameter post-processing on the host: impedance conversions, renormalisation, the CSV and Touchstone 1.1 files
// and the regression comparator, on S matrices as opem_sweep returns them (complex [nports][nports] row-major per
// frequency, the layout of the reference's Result::S Mat).
//
//   Result::S2Z / Z2S / renormalize      src/OpenParEM3D/results.cpp:455-578
//   ResultDatabase::saveCSV              results.cpp:940-1046   (rows: Result::saveCSV, results.cpp:208-232)
//   ResultDatabase::saveTouchstone (1.1) results.cpp:1048-1190
//   Result / ResultDatabase::save_as_test results.cpp:276-341, 1540-1580
//   ResultDatabase::loadCSV, extractS    results.cpp:1582-1640, 343-399
//   TestCase(Database) load / evaluate / show_evaluation   src/OpenParEM3D/process3D.cpp:105-257, 424-457
//
// The files are written with the same iostream formatting as the reference (15 significant digits for data, the
// stream default elsewhere), so a file written here and one written by the reference differ only where the
// numbers do.  Everything is small dense work (nports <= a few tens); none of it touches the GPU.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opem.h"

namespace {

using cplx = std::complex<double>;
using Mat = std::vector<cplx>;  // n x n row-major

thread_local std::string g_err;  // message of the last failed context-free call

Mat load(int n, const double *p) {
  Mat m((size_t)n * n);
  for (size_t i = 0; i < m.size(); i++) m[i] = cplx(p[2 * i], p[2 * i + 1]);
  return m;
}

void store(const Mat &m, double *p) {
  for (size_t i = 0; i < m.size(); i++) {
    p[2 * i] = m[i].real();
    p[2 * i + 1] = m[i].imag();
  }
}

Mat matmul(int n, const Mat &a, const Mat &b) {
  Mat c((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < n; k++) {
      const cplx aik = a[(size_t)i * n + k];
      for (int j = 0; j < n; j++) c[(size_t)i * n + j] += aik * b[(size_t)k * n + j];
    }
  return c;
}

// inverse by Gauss-Jordan elimination with partial pivoting
Mat inverse(int n, Mat a) {
  Mat inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(a[(size_t)i * n + k]) > std::abs(a[(size_t)p * n + k])) p = i;
    if (std::abs(a[(size_t)p * n + k]) == 0.0) throw std::runtime_error("singular matrix");
    if (p != k)
      for (int j = 0; j < n; j++) {
        std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
        std::swap(inv[(size_t)k * n + j], inv[(size_t)p * n + j]);
      }
    const cplx d = 1.0 / a[(size_t)k * n + k];
    for (int j = 0; j < n; j++) {
      a[(size_t)k * n + j] *= d;
      inv[(size_t)k * n + j] *= d;
    }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const cplx f = a[(size_t)i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
        inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j];
      }
    }
  }
  return inv;
}

// k = diag(sqrt(Zo)) (Result::get_k, results.cpp:419-453)
Mat diag_sqrt(int n, const std::vector<cplx> &zo) {
  Mat k((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) k[(size_t)i * n + i] = std::sqrt(zo[i]);
  return k;
}

// Z = k (I + S) (I - S)^-1 k
Mat s2z(int n, const Mat &S, const std::vector<cplx> &zo) {
  Mat ips = S, ims((size_t)n * n);
  for (size_t i = 0; i < S.size(); i++) ims[i] = -S[i];
  for (int i = 0; i < n; i++) {
    ips[(size_t)i * n + i] += 1.0;
    ims[(size_t)i * n + i] += 1.0;
  }
  const Mat k = diag_sqrt(n, zo);
  return matmul(n, matmul(n, matmul(n, k, ips), inverse(n, ims)), k);
}

// S = (Z k^-1 + k)^-1 (Z k^-1 - k)
Mat z2s(int n, const Mat &Z, const std::vector<cplx> &zo) {
  const Mat k = diag_sqrt(n, zo);
  Mat ki((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) ki[(size_t)i * n + i] = 1.0 / k[(size_t)i * n + i];
  const Mat zk = matmul(n, Z, ki);
  Mat p = zk, m = zk;
  for (size_t i = 0; i < zk.size(); i++) {
    p[i] += k[i];
    m[i] -= k[i];
  }
  return matmul(n, inverse(n, p), m);
}

std::vector<cplx> load_vec(int n, const double *p) {
  std::vector<cplx> v(n);
  for (int i = 0; i < n; i++) v[i] = cplx(p[2 * i], p[2 * i + 1]);
  return v;
}

void check_square(int n, const void *a, const void *b, const void *c) {
  if (n < 1 || !a || !b || !c) throw std::invalid_argument("bad arguments");
}

// frequency scale of a unit name; false if the name is not one of Hz, kHz, MHz, GHz
bool unit_scale(const std::string &u, double &scale) {
  if (u == "Hz") scale = 1;
  else if (u == "kHz") scale = 1e-3;
  else if (u == "MHz") scale = 1e-6;
  else if (u == "GHz") scale = 1e-9;
  else return false;
  return true;
}

struct FileMeta {
  std::string version, format, unit;
  double zref = 0;
  std::vector<std::string> nets;
  double scale = 1;
};

FileMeta check_meta(const opem_sparam_file *meta, int nports) {
  if (!meta || !meta->format || !meta->unit) throw std::invalid_argument("bad file description");
  FileMeta m;
  m.version = meta->version ? meta->version : "";
  m.format = meta->format;
  m.unit = meta->unit;
  m.zref = meta->reference_impedance;
  if (m.format != "RI" && m.format != "MA" && m.format != "DB") throw std::invalid_argument("format must be RI, MA or DB");
  if (!unit_scale(m.unit, m.scale)) throw std::invalid_argument("unit must be Hz, kHz, MHz or GHz");
  for (int j = 0; j < nports; j++)
    m.nets.push_back(meta->nets && meta->nets[j] ? std::string(meta->nets[j]) : "net" + std::to_string(j + 1));
  return m;
}

// the two numbers a format writes for one S-parameter
void put_pair(std::ostream &o, const std::string &fmt, cplx s, const char *sep) {
  if (fmt == "RI") o << sep << s.real() << sep << s.imag();
  if (fmt == "MA") o << sep << std::abs(s) << sep << std::arg(s) * 180 / M_PI;
  if (fmt == "DB") o << sep << 20 * std::log10(std::abs(s)) << sep << std::arg(s) * 180 / M_PI;
}

// double_compare of the reference's common code (src/OpenParEMCommon/misc.cpp:76-84)
bool same_frequency(double a, double b, double tol) {
  if (a == b) return true;
  if (a == 0 && std::fabs(b) < tol) return true;
  if (b == 0 && std::fabs(a) < tol) return true;
  if (std::fabs(a) < tol && std::fabs(b) < tol) return true;
  return std::fabs((a - b) / a) < tol;
}

std::vector<std::string> split(const std::string &line, char sep) {
  std::vector<std::string> out;
  std::stringstream ss(line);
  std::string e;
  while (std::getline(ss, e, sep)) out.push_back(e);
  return out;
}

bool parse_double(const std::string &s, double &v) {
  const char *b = s.c_str();
  while (*b == ' ' || *b == '\t') b++;
  if (!*b) return false;
  char *e = nullptr;
  v = std::strtod(b, &e);
  while (e && (*e == ' ' || *e == '\t' || *e == '\r')) e++;
  return e && *e == 0;
}

bool parse_int(const std::string &s, int &v) {
  double d;
  if (!parse_double(s, d) || d != std::floor(d)) return false;
  v = (int)d;
  return true;
}

struct Loaded {  // one frequency of an S-parameter CSV as the comparator reads it: the two columns of each entry
  double freq;
  std::vector<cplx> S;
};

// ResultDatabase::loadCSV: the header gives the frequency unit and port count; every other non-comment line is
// "freq,a11,b11,a12,b12,..." with entry (i, j) stored as a + j b in the order i outer, j inner.
std::vector<Loaded> load_results_csv(const char *path, int &nports) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  std::string line, unit;
  nports = -1;
  std::vector<std::string> lines;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    lines.push_back(line);
    const auto f = split(line, ',');
    for (size_t k = 0; k + 1 < f.size(); k++) {
      if (f[k] == "#frequency unit") unit = f[k + 1];
      if (f[k] == "#number of ports") nports = std::stoi(f[k + 1]);
    }
  }
  double scale = 1;
  if (nports < 1 || !unit_scale(unit, scale)) throw std::runtime_error("results file has no port count or unit");
  std::vector<Loaded> out;
  for (const auto &l : lines) {
    if (l.empty() || l[0] == '#') continue;
    const auto f = split(l, ',');
    if (f.size() != (size_t)(1 + 2 * nports * nports)) throw std::runtime_error("results line of the wrong length");
    Loaded r;
    double v;
    if (!parse_double(f[0], v)) throw std::runtime_error("bad frequency in results file");
    r.freq = v / scale;  // the file's unit back to Hz
    r.S.resize((size_t)nports * nports);
    for (int e = 0; e < nports * nports; e++) {
      double a, b;
      if (!parse_double(f[1 + 2 * e], a) || !parse_double(f[2 + 2 * e], b))
        throw std::runtime_error("bad number in results file");
      r.S[e] = cplx(a, b);
    }
    out.push_back(r);
  }
  return out;
}

struct TestCase {
  std::string name, variable, function;
  double freq = 0, expected = 0, threshold = 0, tolerance = 0, found = 0, error = 0;
  int port_out = 0, port_in = 0;
  bool passed = false, evaluated = false;
};

// TestCaseDatabase::load for S-parameter cases: name,frequency,variable,portOut,portIn,function,value[,tolerance]
std::vector<TestCase> load_cases(const char *path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  std::vector<TestCase> cases;
  std::string line;
  int ln = 0;
  while (std::getline(in, line)) {
    ln++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    line = line.substr(0, line.find("//"));
    const auto f = split(line, ',');
    TestCase t;
    bool ok = f.size() >= 7;
    if (ok) {
      t.name = f[0];
      ok = parse_double(f[1], t.freq) && (f[2] == "magS(dB)" || f[2] == "argS(deg)") && parse_int(f[3], t.port_out) &&
           parse_int(f[4], t.port_in) && (f[5] == "equal" || f[5] == "lessthan" || f[5] == "greaterthan");
      t.variable = f[2];
      t.function = f[5];
    }
    if (ok && t.function == "equal")
      ok = f.size() == 8 && parse_double(f[6], t.expected) && parse_double(f[7], t.tolerance);
    else if (ok)
      ok = f.size() == 7 && parse_double(f[6], t.threshold);
    if (!ok) throw std::runtime_error("bad test case on line " + std::to_string(ln) + " of " + path);
    cases.push_back(t);
  }
  return cases;
}

// TestCase::evaluate for S-parameter cases: the comparator reads the first column of an entry as dB and the second
// as degrees (files written in DB format); entry (portIn, portOut) of the loaded matrix
void evaluate(TestCase &t, const std::vector<Loaded> &res, int nports) {
  t.passed = false;
  t.evaluated = false;
  const Loaded *r = nullptr;
  for (const auto &x : res)
    if (same_frequency(x.freq, t.freq, 1e-12)) {
      r = &x;
      break;
    }
  if (r && t.port_in >= 1 && t.port_in <= nports && t.port_out >= 1 && t.port_out <= nports) {
    const cplx v = r->S[(size_t)(t.port_in - 1) * nports + (t.port_out - 1)];
    if (t.variable == "magS(dB)") {
      t.found = v.real();
      if (t.function == "equal") {
        t.error = t.expected == 0 ? std::fabs(t.found) : std::fabs((t.found - t.expected) / t.expected);
        t.passed = t.error <= t.tolerance;
      }
      if (t.function == "lessthan") t.passed = t.found <= t.threshold;
      if (t.function == "greaterthan") t.passed = t.found >= t.threshold;
    } else {
      t.found = v.imag();
      if (t.function == "equal") {
        if (t.expected == 0) {
          t.error = std::fabs(t.found);
        } else {  // a wrap of +-360 degrees is no error
          const double e1 = std::fabs((t.found - t.expected) / t.expected);
          const double e2 = std::fabs((t.found + 360 - t.expected) / t.expected);
          const double e3 = std::fabs((t.found - 360 - t.expected) / t.expected);
          t.error = std::min(std::min(e1, e2), e3);
        }
        t.passed = t.error <= t.tolerance;
      } else {
        while (t.found >= 180) t.found -= 360;
        while (t.found <= -180) t.found += 360;
        t.passed = t.function == "lessthan" ? t.found <= t.threshold : t.found >= t.threshold;
      }
    }
  }
  t.evaluated = true;
}

double sort_key(const TestCase &t) { return t.function == "equal" ? t.error : t.found; }

template <class F>
int guarded_free(F f) {
  try {
    f();
    g_err.clear();
    return 0;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return -2;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  }
}

}  // namespace

extern "C" {

const char *opem_results_error(void) { return g_err.c_str(); }

int opem_s2z(int nports, const double *S, const double *Zo, double *Z) {
  return guarded_free([&] {
    check_square(nports, S, Zo, Z);
    store(s2z(nports, load(nports, S), load_vec(nports, Zo)), Z);
  });
}

int opem_z2s(int nports, const double *Z, const double *Zo, double *S) {
  return guarded_free([&] {
    check_square(nports, Z, Zo, S);
    store(z2s(nports, load(nports, Z), load_vec(nports, Zo)), S);
  });
}

int opem_renormalize(int nports, const double *S, const double *Zo, double zref_re, double zref_im, double *S_out) {
  return guarded_free([&] {
    check_square(nports, S, Zo, S_out);
    if (!(zref_re != 0 || zref_im != 0)) throw std::invalid_argument("reference impedance must be nonzero");
    const Mat Z = s2z(nports, load(nports, S), load_vec(nports, Zo));
    store(z2s(nports, Z, std::vector<cplx>(nports, cplx(zref_re, zref_im))), S_out);
  });
}

int opem_write_csv(const char *path, const opem_sparam_file *meta, int nports, int64_t nfreq, const double *freqs,
                   const double *S) {
  return guarded_free([&] {
    if (!path || nports < 1 || nfreq < 0 || (nfreq > 0 && (!freqs || !S))) throw std::invalid_argument("bad arguments");
    const FileMeta m = check_meta(meta, nports);
    std::ofstream o(path);
    if (!o) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
    if (!m.version.empty()) o << "#OpenParEM3D " << m.version << std::endl;
    o << "#Touchstone format," << m.format << std::endl;
    o << "#frequency unit," << m.unit << std::endl;
    o << "#number of frequencies," << nfreq << std::endl;
    o << "#number of ports," << nports << std::endl;
    for (int j = 0; j < nports; j++) {
      o << "#S-port " << j + 1 << "," << m.nets[j] << ",";
      if (m.zref == 0) o << "not renormalized";
      else o << m.zref;
      o << std::endl;
    }
    o << "#Frequency(" << m.unit << ")";
    const char *a = m.format == "RI" ? "Re" : m.format == "MA" ? "mag" : "dB";
    const char *b = m.format == "RI" ? "Im" : "deg";
    for (int col = 0; col < nports; col++)
      for (int row = 0; row < nports; row++)
        o << "," << a << "(S(" << row + 1 << ";" << col + 1 << "))," << b << "(S(" << row + 1 << ";" << col + 1 << "))";
    o << std::endl;
    for (int64_t f = 0; f < nfreq; f++) {
      o << std::setprecision(15) << freqs[f] * m.scale;
      const double *Sf = S + (size_t)f * nports * nports * 2;
      for (int i = 0; i < nports; i++)
        for (int j = 0; j < nports; j++) {
          const size_t e = (size_t)i * nports + j;
          put_pair(o, m.format, cplx(Sf[2 * e], Sf[2 * e + 1]), ",");
        }
      o << std::setprecision(6) << std::endl;
    }
    if (!o) throw std::runtime_error(std::string("write failed: ") + path);
  });
}

int opem_write_touchstone(const char *path, const opem_sparam_file *meta, int nports, int64_t nfreq,
                          const double *freqs, const double *S) {
  return guarded_free([&] {
    if (!path || nports < 1 || nfreq < 0 || (nfreq > 0 && (!freqs || !S))) throw std::invalid_argument("bad arguments");
    const FileMeta m = check_meta(meta, nports);
    // a frequency-dependent reference impedance has no Touchstone 1.1 form (results.cpp:1056-1063)
    if (m.zref == 0) throw std::invalid_argument("Touchstone output needs renormalized data (reference_impedance > 0)");
    std::ofstream o(path);
    if (!o) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
    if (!m.version.empty()) o << "! OpenParEM3D " << m.version << std::endl;
    o << "! " << nports << "-port S-parameter data" << std::endl;
    for (int i = 0; i < nports; i++) o << "! S-port " << i + 1 << " " << m.nets[i] << std::endl;
    o << "# " << m.unit << " S " << m.format << " R " << m.zref << std::endl;
    const char *a = m.format == "RI" ? "ReS" : m.format == "MA" ? "magS" : "dBS";
    const char *b = m.format == "RI" ? "ImS" : "angS";
    if (nports < 3) {
      o << "!freq";
      for (int col = 0; col < nports; col++)
        for (int row = 0; row < nports; row++)
          o << " " << a << row + 1 << col + 1 << " " << b << row + 1 << col + 1;
      o << std::endl;
    }
    for (int64_t f = 0; f < nfreq; f++) {
      o << std::setprecision(15) << freqs[f] * m.scale;
      const double *Sf = S + (size_t)f * nports * nports * 2;
      for (int i = 1; i <= nports; i++) {
        int count = 0;
        bool row_printed = false;
        for (int j = 1; j <= nports; j++) {
          // one- and two-port files list the matrix column by column (N11 N21 N12 N22), larger ones row by row
          const size_t e = nports < 3 ? (size_t)(j - 1) * nports + (i - 1) : (size_t)(i - 1) * nports + (j - 1);
          put_pair(o, m.format, cplx(Sf[2 * e], Sf[2 * e + 1]), " ");
          count++;
          if (nports > 2 && (count == nports || count == 4)) {
            if (!row_printed) {
              o << " !row " << i;
              row_printed = true;
            }
            o << std::endl;
            count = 0;
          }
        }
        if (nports > 2 && !row_printed) o << " !row " << i;
        if (nports > 2 && count > 0) o << std::endl;
      }
      if (nports <= 2) o << std::endl;
    }
    if (!o) throw std::runtime_error(std::string("write failed: ") + path);
  });
}

int opem_save_as_test(const char *path, const char *casename, int nports, int64_t nfreq, const double *freqs,
                      const double *S, int *casenumber) {
  return guarded_free([&] {
    if (!path || !casename || nports < 1 || nfreq < 0 || (nfreq > 0 && (!freqs || !S)) || !casenumber)
      throw std::invalid_argument("bad arguments");
    const double mag_limit_db = -150, equal_mag_limit = 1e-12, arg_limit_deg = 0.1, equal_arg_limit = 1e-12;
    std::ofstream o(path, std::ios::app);
    if (!o) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
    o << "# ResultDatabase::save_as_test" << std::endl;
    for (int64_t f = 0; f < nfreq; f++) {
      const double *Sf = S + (size_t)f * nports * nports * 2;
      for (int i = 0; i < nports; i++)
        for (int j = 0; j < nports; j++) {
          const size_t e = (size_t)i * nports + j;
          const cplx v(Sf[2 * e], Sf[2 * e + 1]);
          const double db = 20 * std::log10(std::abs(v)), deg = std::arg(v) * 180 / M_PI;
          o << casename << "_" << f + 1 << "_" << (*casenumber)++ << "_result" << ",";
          o << std::setprecision(15) << freqs[f] << ",magS(dB)," << j + 1 << "," << i + 1 << ",";
          if (db > mag_limit_db) o << "equal," << std::setprecision(15) << db << "," << equal_mag_limit << std::endl;
          else o << "lessthan," << mag_limit_db << std::endl;
          if (db > mag_limit_db) {
            o << casename << "_" << f + 1 << "_" << (*casenumber)++ << "_result" << ",";
            o << std::setprecision(15) << freqs[f] << ",argS(deg)," << j + 1 << "," << i + 1 << ",";
            if (std::fabs(deg) > arg_limit_deg) o << "equal," << std::setprecision(15) << deg << "," << equal_arg_limit << std::endl;
            else o << "lessthan," << arg_limit_deg << std::endl;
          }
        }
    }
    if (!o) throw std::runtime_error(std::string("write failed: ") + path);
  });
}

int opem_process_tests(const char *cases_path, const char *results_csv, const char *report_path, int *ncases,
                       int *nevaluated, int *npassed) {
  return guarded_free([&] {
    if (!cases_path || !results_csv || !report_path) throw std::invalid_argument("bad arguments");
    std::vector<TestCase> cases = load_cases(cases_path);
    int nports = 0;
    const std::vector<Loaded> res = load_results_csv(results_csv, nports);
    for (auto &t : cases) evaluate(t, res, nports);
    // largest |error| (equal) or |found value| (thresholds) first, as TestCaseDatabase::sort orders the report
    std::vector<size_t> order(cases.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    for (size_t i = 0; i + 1 < order.size(); i++)
      for (size_t j = i + 1; j < order.size(); j++)
        if (std::fabs(sort_key(cases[order[i]])) < std::fabs(sort_key(cases[order[j]]))) std::swap(order[i], order[j]);
    std::ofstream o(report_path);
    if (!o) throw std::runtime_error(std::string("cannot open ") + report_path + " for writing");
    o << "# name,status,frequency,portOut,portIn,testVariable,testFunction,[equal:expectedValue,]foundValue,"
      << "[equal:error,tolerance | lessthan:threshold | greaterthan:threshold" << std::endl;
    int ev = 0, pa = 0;
    for (size_t k : order) {
      const TestCase &t = cases[k];
      ev += t.evaluated;
      pa += t.passed;
      if (!t.evaluated) continue;
      o << t.name << "," << (t.passed ? "pass" : "FAIL") << "," << std::setprecision(15) << t.freq << "," << t.port_out
        << "," << t.port_in << "," << t.variable << "," << t.function << ",";
      if (t.function == "equal") o << std::setprecision(15) << t.expected << ",";
      o << std::setprecision(15) << t.found << ",";
      if (t.function == "equal") o << std::setprecision(15) << t.error << "," << t.tolerance << ",";
      else o << t.threshold << ",";
      o << std::endl;
    }
    if (ncases) *ncases = (int)cases.size();
    if (nevaluated) *nevaluated = ev;
    if (npassed) *npassed = pa;
  });
}

}  // extern "C"
