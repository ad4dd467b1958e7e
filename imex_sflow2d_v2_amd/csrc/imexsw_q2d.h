// .q_2d state files: the reference's formatted restart/output format
// (output_solution src/inpout_2d.f90:5766-5891, read_solution :5432-5453).
// One record per cell, j fastest within k: x_comp(j), y_comp(k), q(1..nv),
// each field Fortran E20.12; a record with one blank after every row of
// cells, two at the end.  Host code only (no device work).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace imexsw {

// Fortran E20.12 output editing: 0.dddddddddddd with the exponent as E+dd,
// or as +ddd without the letter when it needs three digits; right-justified
// in 20 columns.  The digits are the correctly rounded 12 significant
// digits of the binary value (printf %.11e gives d.ddddddddddd, which is the
// same digit string with the exponent one smaller).
// Not-a-number and infinities are written as gfortran writes them in a
// 20-column E field: NaN, Infinity, -Infinity, right-justified.
inline void fortran_e20_12(double v, char out[21]) {
  if (!std::isfinite(v)) {
    std::snprintf(out, 21, "%20s", std::isnan(v) ? "NaN" : (v < 0.0 ? "-Infinity" : "Infinity"));
    return;
  }
  char m[40] = {0};
  std::snprintf(m, sizeof m, "%.11e", std::fabs(v));
  // m = "d.ddddddddddde[+-]XX[X]"
  char digits[13];
  digits[0] = m[0];
  std::memcpy(digits + 1, m + 2, 11);
  digits[12] = 0;
  int e = std::atoi(m + 14);
  if (v != 0.0) e += 1;  // 0.d... form; zero keeps exponent 0
  char body[32];
  const char* sign = std::signbit(v) && v != 0.0 ? "-" : "";
  const int ae = e < 0 ? -e : e;
  if (ae <= 99)
    std::snprintf(body, sizeof body, "%s0.%sE%c%02d", sign, digits, e < 0 ? '-' : '+', ae);
  else
    std::snprintf(body, sizeof body, "%s0.%s%c%03d", sign, digits, e < 0 ? '-' : '+', ae);
  std::snprintf(out, 21, "%20s", body);
}

// Fortran E input editing of one fixed-width field: blanks are ignored, an
// empty field is zero, the exponent letter (E, e, D, d) may be missing when
// the exponent carries a sign ("0.1-100").  Returns false for a malformed
// field.
inline bool fortran_e_field(const char* f, int w, double* v) {
  std::string s;
  for (int i = 0; i < w && f[i]; i++)
    if (f[i] != ' ' && f[i] != '\t' && f[i] != '\r' && f[i] != '\n') s += f[i];
  if (s.empty()) {
    *v = 0.0;
    return true;
  }
  for (char& ch : s)
    if (ch == 'd' || ch == 'D') ch = 'E';
  // a sign after a digit or the decimal point starts an exponent
  for (size_t i = 1; i < s.size(); i++)
    if ((s[i] == '+' || s[i] == '-') && (std::isdigit((unsigned char)s[i - 1]) || s[i - 1] == '.')) {
      s.insert(i, 1, 'E');
      break;
    }
  char* end = nullptr;
  *v = std::strtod(s.c_str(), &end);
  return end && *end == 0;
}

// Reads one record (line) into `line`; false at end of file.
inline bool read_record(FILE* fp, std::string& line) {
  line.clear();
  int ch;
  bool any = false;
  while ((ch = std::fgetc(fp)) != EOF) {
    any = true;
    if (ch == '\n') return true;
    line += (char)ch;
  }
  return any;
}

}  // namespace imexsw
