// LDS layout of the in-tile wavefront of the tiled inside pass (finalising
// workgroups), shared by the kernels (inside.hip) and a host-side checker
// (tests/wave_layout_check.cpp).  Plain C++: no HIP.
//
// The wavefront of an off-diagonal T x T tile (I, J) finalises cell (a, b) at
// step own = b + T-1 - a.  At step st a cell adds two split-point terms whose
// newest operand was finalised at step st-1:
//   column term  Ld[a][ap-1] * cS[ap][b],   ap = b + T - st   (S below, same column)
//   row term     cL[a][bp]   * Sd[bp+1][b], bp = st - T + a   (L to the left, same row)
// with Ld / Sd the diagonal tiles (I, I) / (J, J) and cS / cL the tile itself.
// The column term is needed for a < ap <= T (ap == T: the split point on the
// tile's bottom border, S[(I+1)T, y], held as row T of cS), the row term for
// 0 <= bp < b.  Every other (lane, step) read is of one of two kinds:
//   * zero: ap > T or bp < 0 -- the term must be zero, and the layout puts a
//     zero cell under one operand and a finite cell under the other;
//   * dead: st > own -- the cell is final, its accumulator is no longer used,
//     and the read only has to stay inside the array.
// So a step reads four cells at a per-lane base plus a compile-time offset of
// st, with no index arithmetic and no masks.  The regions:
//
//   [Ld pad T][Ld T x RS][Ld pad T] [Sd T x RS] [cP T x RS] | [cF T x RS][cS 2T x RS][cL T x RL + T]
//   <------------------ staged before the products -----------> <------ after them (a union) ----->
//
// RS = T + 1 (row stride; column T is a pad), RL = 2T + 1: cL row a holds T zero
// cells, then its T cells at columns 0..T-1, then a pad.  Zero regions: Ld's pads
// and pad column; cS rows T+1 .. 2T-1 and its pad column; cL's zero columns, pad
// column and tail.  Sd's rows above row 0 (read with a zero cL cell) fall in the
// Ld region, which is finite everywhere; cS's rows above row 0 (dead reads) fall
// in cF; Sd's rows below row T-1 (dead) in cP.
#pragma once

namespace messi {
namespace wave {

template <int T> struct Layout {
  static constexpr int RS = T + 1;
  static constexpr int RL = 2 * T + 1;
  // element offsets (8-byte cells) into one flat array
  static constexpr int LD_PAD0 = 0;                  // T zero cells
  static constexpr int LD = LD_PAD0 + T;             // Ld[a][c] at LD + a*RS + c
  static constexpr int LD_PAD1 = LD + T * RS;        // T zero cells
  static constexpr int SD = LD_PAD1 + T;             // Sd[r][b] at SD + r*RS + b
  static constexpr int CP = SD + T * RS;             // cP[a][b] at CP + a*RS + b
  static constexpr int STAGED = CP + T * RS;         // end of the part staged before the products
  static constexpr int CF = STAGED;                  // cF[a][b] at CF + a*RS + b
  static constexpr int CS = CF + T * RS;             // cS[r][b] at CS + r*RS + b, r in 0 .. 2T-1
  static constexpr int CL = CS + 2 * T * RS;         // cL[a][c] at CL + a*RL + T + c
  static constexpr int END = CL + T * RL + T;
  static constexpr int WAVE = END - STAGED;          // cells of the part after the products

  static constexpr int ld(int a, int c) { return LD + a * RS + c; }
  static constexpr int sd(int r, int b) { return SD + r * RS + b; }
  static constexpr int cp(int a, int b) { return CP + a * RS + b; }
  static constexpr int cf(int a, int b) { return CF + a * RS + b; }
  static constexpr int cs(int r, int b) { return CS + r * RS + b; }
  static constexpr int cl(int a, int c) { return CL + a * RL + T + c; }

  // The four reads of lane (a, b) at step st, as per-lane bases plus offsets
  // that are constants once st is (the loop over st is unrolled).
  static constexpr int LAST = 2 * T - 2;  // the last step
  static constexpr int base_col_l(int a, int b) { return ld(a, b + T - 1 - LAST); }
  static constexpr int off_col_l(int st) { return LAST - st; }
  static constexpr int base_col_s(int, int b) { return cs(b + T - LAST, b); }
  static constexpr int off_col_s(int st) { return (LAST - st) * RS; }
  static constexpr int base_row_l(int a, int) { return cl(a, a - T); }
  static constexpr int off_row_l(int st) { return st; }
  static constexpr int base_row_s(int a, int b) { return sd(a + 1 - T, b); }
  static constexpr int off_row_s(int st) { return st * RS; }

  // Cells set to zero before the wavefront: k = 0 .. N_ZERO_STAGED-1 in the staged
  // part (Ld's two pads and its pad column), k = 0 .. N_ZERO_WAVE-1 in the part
  // after the products (cS rows T+1 .. 2T-1, the pad cell of cS row T, cL's zero
  // columns, pad column and tail).  cS row T (columns 0..T-1) receives the
  // bottom-border S row.
  static constexpr int N_ZERO_STAGED = 3 * T;
  static constexpr int zero_staged(int k) {
    return k < T ? LD_PAD0 + k : (k < 2 * T ? ld(k - T, T) : LD_PAD1 + (k - 2 * T));
  }
  static constexpr int NZ_CS = (T - 1) * RS + 1, NZ_CL = T * T + 2 * T;
  static constexpr int N_ZERO_WAVE = NZ_CS + NZ_CL;
  static constexpr int zero_wave(int k) {
    return k < NZ_CS - 1 ? cs(T + 1, 0) + k
         : k == NZ_CS - 1 ? cs(T, T)
         : (k - NZ_CS) < T * T ? cl((k - NZ_CS) / T, (k - NZ_CS) % T - T)
         : (k - NZ_CS - T * T) < T ? cl(k - NZ_CS - T * T, T)
         : CL + T * RL + (k - NZ_CS - T * T - T);
  }
};

}  // namespace wave
}  // namespace messi
