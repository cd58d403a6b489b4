// cs_fused_weights.h -- the reference CLIs' weight set-up, as data.
//
// The reference's drivers turn user weights into pre-fused weights for their
// "*1r" shapes before launching (temporal fusion by weight composition):
//   1d1r: 3 weights -> 7 (src/1d/main.cu:133-140), 1d2r: 5 -> 9 (:141-151),
//   2-D:  9 weights of a 3x3 stencil -> 49 of a 7x7 (src/2d/main.cu:233-281),
//   3-D:  fixed defaults i/100 (box, 196) and 1 (star, 44) (src/3d/main.cu:526-534).
// The 2-D set-up is a hand-expanded table of the cube of the 3x3 stencil in which
// nine entries (18, 19, 30-35, 41) do not equal the exact composition unless all
// weights are equal.  A drop-in must reproduce the reference's numbers, so the
// table is kept here as data -- for every output coefficient the list of terms
// coef * p[a] * p[b] * p[c], evaluated ((coef * p[a]) * p[b]) * p[c] and summed left
// to right as the reference's C expressions are; coef 0 marks the one term the
// reference writes as pow(p[a], 3).  fused_2d_exact() is the exact composition.
#pragma once
#include <cmath>

namespace cs {

struct FusedTerm {
  int coef, a, b, c;
};
struct FusedEntry {
  int n;
  FusedTerm t[9];
};

// Coefficient i of the reference's 49-entry table (row-major 7x7, dy outer).
static const FusedEntry kRef2DTable[49] = {
    /*  0 */ {1, {{1, 0, 0, 0}}},
    /*  1 */ {1, {{3, 0, 0, 1}}},
    /*  2 */ {2, {{3, 0, 0, 2}, {3, 0, 1, 1}}},
    /*  3 */ {2, {{6, 0, 1, 2}, {1, 1, 1, 1}}},
    /*  4 */ {2, {{3, 0, 2, 2}, {3, 1, 1, 2}}},
    /*  5 */ {1, {{3, 1, 2, 2}}},
    /*  6 */ {1, {{1, 2, 2, 2}}},
    /*  7 */ {1, {{3, 0, 0, 3}}},
    /*  8 */ {2, {{3, 0, 0, 4}, {6, 0, 1, 3}}},
    /*  9 */ {4, {{3, 0, 0, 5}, {6, 0, 1, 4}, {6, 0, 2, 3}, {3, 1, 1, 3}}},
    /* 10 */ {4, {{6, 0, 1, 5}, {6, 0, 2, 4}, {3, 1, 1, 4}, {6, 1, 2, 3}}},
    /* 11 */ {4, {{6, 0, 2, 5}, {3, 1, 1, 5}, {6, 1, 2, 4}, {3, 2, 2, 3}}},
    /* 12 */ {2, {{6, 1, 2, 5}, {3, 2, 2, 4}}},
    /* 13 */ {1, {{3, 2, 2, 5}}},
    /* 14 */ {2, {{3, 0, 0, 6}, {3, 0, 3, 3}}},
    /* 15 */ {4, {{3, 0, 0, 7}, {6, 0, 1, 6}, {6, 0, 3, 4}, {3, 1, 3, 3}}},
    /* 16 */ {8, {{3, 0, 0, 8}, {6, 0, 1, 7}, {6, 0, 2, 6}, {6, 0, 3, 5}, {3, 0, 4, 4}, {3, 1, 1, 6}, {6, 1, 3, 4}, {3, 2, 3, 3}}},
    /* 17 */ {8, {{6, 0, 1, 8}, {6, 0, 2, 7}, {6, 0, 4, 5}, {3, 1, 1, 7}, {6, 1, 2, 6}, {6, 1, 3, 5}, {3, 1, 4, 4}, {6, 2, 3, 4}}},
    /* 18 */ {8, {{6, 0, 2, 8}, {3, 0, 0, 5}, {3, 1, 1, 8}, {6, 1, 2, 7}, {6, 1, 4, 5}, {3, 2, 2, 6}, {6, 2, 3, 5}, {3, 2, 4, 4}}},
    /* 19 */ {4, {{6, 1, 2, 8}, {3, 1, 1, 5}, {3, 2, 2, 7}, {6, 2, 4, 5}}},
    /* 20 */ {2, {{3, 2, 2, 8}, {3, 2, 5, 5}}},
    /* 21 */ {2, {{6, 0, 3, 6}, {1, 3, 3, 3}}},
    /* 22 */ {4, {{6, 0, 3, 7}, {6, 0, 4, 6}, {6, 1, 3, 6}, {3, 3, 3, 4}}},
    /* 23 */ {8, {{6, 0, 3, 8}, {6, 0, 4, 7}, {6, 0, 5, 6}, {6, 1, 3, 7}, {6, 1, 4, 6}, {6, 2, 3, 6}, {3, 3, 3, 5}, {3, 3, 4, 4}}},
    /* 24 */ {9, {{6, 0, 4, 8}, {6, 0, 5, 7}, {6, 1, 3, 8}, {6, 1, 4, 7}, {6, 1, 5, 6}, {6, 2, 3, 7}, {6, 2, 4, 6}, {6, 3, 4, 5}, {0, 4, 4, 4}}},
    /* 25 */ {8, {{6, 0, 5, 8}, {6, 1, 4, 8}, {6, 1, 5, 7}, {6, 2, 3, 8}, {6, 2, 4, 7}, {6, 2, 5, 6}, {3, 3, 5, 5}, {3, 4, 4, 5}}},
    /* 26 */ {4, {{6, 1, 5, 8}, {6, 2, 4, 8}, {6, 2, 5, 7}, {3, 4, 5, 5}}},
    /* 27 */ {2, {{6, 2, 5, 8}, {1, 5, 5, 5}}},
    /* 28 */ {2, {{3, 0, 6, 6}, {3, 3, 3, 6}}},
    /* 29 */ {4, {{6, 0, 6, 7}, {3, 1, 6, 6}, {3, 3, 3, 7}, {6, 3, 4, 6}}},
    /* 30 */ {8, {{6, 0, 6, 8}, {3, 0, 7, 7}, {6, 1, 6, 7}, {3, 2, 6, 6}, {3, 3, 3, 8}, {6, 3, 4, 7}, {6, 3, 5, 6}, {3, 4, 6, 6}}},
    /* 31 */ {8, {{6, 0, 7, 8}, {6, 1, 6, 8}, {3, 1, 7, 7}, {6, 2, 6, 7}, {6, 3, 4, 8}, {6, 3, 5, 7}, {3, 4, 7, 7}, {6, 4, 5, 6}}},
    /* 32 */ {8, {{3, 0, 8, 8}, {6, 1, 7, 8}, {6, 2, 6, 8}, {3, 2, 7, 7}, {6, 3, 5, 8}, {3, 4, 8, 8}, {6, 4, 5, 7}, {3, 5, 6, 6}}},
    /* 33 */ {4, {{3, 1, 8, 8}, {6, 2, 7, 8}, {6, 4, 5, 8}, {3, 5, 7, 7}}},
    /* 34 */ {2, {{3, 2, 8, 8}, {3, 5, 8, 8}}},
    /* 35 */ {1, {{3, 3, 3, 6}}},
    /* 36 */ {2, {{6, 3, 6, 7}, {3, 4, 6, 6}}},
    /* 37 */ {4, {{6, 3, 6, 8}, {3, 3, 7, 7}, {6, 4, 6, 7}, {3, 5, 6, 6}}},
    /* 38 */ {4, {{6, 3, 7, 8}, {6, 4, 6, 8}, {3, 4, 7, 7}, {6, 5, 6, 7}}},
    /* 39 */ {4, {{3, 3, 8, 8}, {6, 4, 7, 8}, {6, 5, 6, 8}, {3, 5, 7, 7}}},
    /* 40 */ {2, {{3, 4, 8, 8}, {6, 5, 7, 8}}},
    /* 41 */ {1, {{3, 5, 5, 8}}},
    /* 42 */ {1, {{1, 6, 6, 6}}},
    /* 43 */ {1, {{3, 6, 6, 7}}},
    /* 44 */ {2, {{3, 6, 6, 8}, {3, 6, 7, 7}}},
    /* 45 */ {2, {{6, 6, 7, 8}, {1, 7, 7, 7}}},
    /* 46 */ {2, {{3, 6, 8, 8}, {3, 7, 7, 8}}},
    /* 47 */ {1, {{3, 7, 8, 8}}},
    /* 48 */ {1, {{1, 8, 8, 8}}},
};

inline void fused_2d_reference(const double p[9], double out[49]) {
  for (int i = 0; i < 49; ++i) {
    double s = 0.0;
    for (int k = 0; k < kRef2DTable[i].n; ++k) {
      const FusedTerm &t = kRef2DTable[i].t[k];
      const double v = t.coef == 0 ? std::pow(p[t.a], 3) : ((t.coef * p[t.a]) * p[t.b]) * p[t.c];
      s = k == 0 ? v : s + v;
    }
    out[i] = s;
  }
}

// Exact 3-fold self-composition of a 3x3 stencil (row-major) into a 7x7 one.
inline void fused_2d_exact(const double p[9], double out[49]) {
  double two[25] = {0.0};
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b) two[(a / 3 + b / 3) * 5 + (a % 3 + b % 3)] += p[a] * p[b];
  for (int i = 0; i < 49; ++i) out[i] = 0.0;
  for (int a = 0; a < 25; ++a)
    for (int b = 0; b < 9; ++b) out[(a / 5 + b / 3) * 7 + (a % 5 + b % 3)] += two[a] * p[b];
}

inline void fused_1d1r(const double v[3], double out[7]) {
  out[0] = v[0] * v[0] * v[0];
  out[1] = 3 * v[0] * v[0] * v[1];
  out[2] = 3 * v[0] * v[0] * v[2] + 3 * v[0] * v[1] * v[1];
  out[3] = 6 * v[0] * v[1] * v[2] + v[1] * v[1] * v[1];
  out[4] = 3 * v[0] * v[2] * v[2] + 3 * v[1] * v[1] * v[2];
  out[5] = 3 * v[1] * v[2] * v[2];
  out[6] = v[2] * v[2] * v[2];
}

inline void fused_1d2r(const double v[5], double out[9]) {
  out[0] = v[0] * v[0];
  out[1] = 2 * v[0] * v[1];
  out[2] = 2 * v[0] * v[2] + v[1] * v[1];
  out[3] = 2 * v[0] * v[3] + 2 * v[1] * v[2];
  out[4] = 2 * v[0] * v[4] + 2 * v[1] * v[3] + v[2] * v[2];
  out[5] = 2 * v[1] * v[4] + 2 * v[2] * v[3];
  out[6] = 2 * v[2] * v[4] + v[3] * v[3];
  out[7] = 2 * v[3] * v[4];
  out[8] = v[4] * v[4];
}

}  // namespace cs
