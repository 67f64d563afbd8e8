// d3q19.h -- D3Q19 velocity set and mode transforms as compile-time tables.
//
// Velocities and weights follow lb_d3q19.h:24-36 (the opposite of p >= 1 is 19 - p).
// The mode matrix rows are {1, c, cc - delta/3, chi1, chi1 c, chi2, chi2 c,
// chi3} (lb_d3q19.c:107-153), the normalisers na = 1/sum_p w ma^2
// (lb_d3q19.c:72-80) and the inverse mi = w na ma^T (lb_data.c:644).  Every
// entry is an exact rational; it is rounded once to double by an IEEE
// division of two exactly representable integers, which is the value the
// reference's unrolled constants (collision.c:1966-1988) hold.
//
// The f -> mode table reproduces one quirk of the reference's unrolled
// transform: ghost mode 13 omits the f[4] term (collision.c:2278 multiplies
// by zero where chi1*cz = -1).  Under the default M10 scheme the ghost
// modes relax to zero and this cannot be seen; under BGK and TRT it can.

#ifndef LDW_D3Q19_H
#define LDW_D3Q19_H

namespace ldw {

constexpr int NVEL = 19;

struct Cv { int v[NVEL][3]; };
constexpr Cv CV = {{
  { 0,  0,  0},
  { 1,  1,  0}, { 1,  0,  1}, { 1,  0,  0},
  { 1,  0, -1}, { 1, -1,  0}, { 0,  1,  1},
  { 0,  1,  0}, { 0,  1, -1}, { 0,  0,  1},
  { 0,  0, -1}, { 0, -1,  1}, { 0, -1,  0},
  { 0, -1, -1}, {-1,  1,  0}, {-1,  0,  1},
  {-1,  0,  0}, {-1,  0, -1}, {-1, -1,  0}}};

// weights as numerators over 36
struct W36 { int v[NVEL]; };
constexpr W36 WV36 = {{12, 1, 1, 2, 1, 1, 1, 2, 1, 2, 2, 1, 2, 1, 1, 1, 2, 1, 1}};

constexpr double weight(int p) { return double(WV36.v[p]) / 36.0; }

// numerator of ma[m][p]; the denominator is 3 for rows 4, 7, 9, else 1
constexpr int ma_num(int m, int p) {
  const int cx = CV.v[p][0], cy = CV.v[p][1], cz = CV.v[p][2];
  const int c2 = cx*cx + cy*cy + cz*cz;
  const int chi1 = (2*c2 - 3)*(3*cz*cz - c2);
  const int chi2 = (2*c2 - 3)*(cy*cy - cx*cx);
  const int chi3 = 3*c2*c2 - 6*c2 + 1;
  switch (m) {
  case 0: return 1;
  case 1: return cx;
  case 2: return cy;
  case 3: return cz;
  case 4: return 3*cx*cx - 1;
  case 5: return cx*cy;
  case 6: return cx*cz;
  case 7: return 3*cy*cy - 1;
  case 8: return cy*cz;
  case 9: return 3*cz*cz - 1;
  case 10: return chi1;
  case 11: return chi1*cx;
  case 12: return chi1*cy;
  case 13: return chi1*cz;
  case 14: return chi2;
  case 15: return chi2*cx;
  case 16: return chi2*cy;
  case 17: return chi2*cz;
  default: return chi3;
  }
}

constexpr int ma_den(int m) { return (m == 4 || m == 7 || m == 9) ? 3 : 1; }

// S_m = sum_p w36_p num_mp^2, so that na_m = 36 den_m^2 / S_m
constexpr long long ma_norm(int m) {
  long long s = 0;
  for (int p = 0; p < NVEL; p++) s += (long long) WV36.v[p]*ma_num(m, p)*ma_num(m, p);
  return s;
}

struct Table { double v[NVEL][NVEL]; };

// f -> mode coefficients as executed by the reference (with the mode-13 quirk)
constexpr Table make_ma_exec() {
  Table t{};
  for (int m = 0; m < NVEL; m++)
    for (int p = 0; p < NVEL; p++)
      t.v[m][p] = double(ma_num(m, p)) / double(ma_den(m));
  t.v[13][4] = 0.0;
  return t;
}

// mode -> f: mi[p][m] = w_p na_m ma_mp = w36_p den_m num_mp / S_m
constexpr Table make_mi() {
  Table t{};
  for (int p = 0; p < NVEL; p++)
    for (int m = 0; m < NVEL; m++)
      t.v[p][m] = double((long long) WV36.v[p]*ma_den(m)*ma_num(m, p)) / double(ma_norm(m));
  return t;
}

constexpr Table MA = make_ma_exec();
constexpr Table MI = make_mi();

}  // namespace ldw

#endif
