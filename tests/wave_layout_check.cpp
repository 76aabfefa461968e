// Host check of the in-tile wavefront's LDS layout (messi_amd/csrc/wave_layout.h),
// the exact index functions the kernel uses.  For every lane (a, b) and step st of
// an off-diagonal T x T tile it classifies both split-point terms and checks:
//   * every read lies in [0, END);
//   * a needed term reads the intended cells, and each was written (finalised at
//     an earlier step, or staged before the wavefront);
//   * a zero term reads a zero cell for one operand and a cell holding a finite
//     value for the other (never a cell not yet written);
//   * a dead term (st > own) only has to stay in bounds;
//   * the cells the finalisations write are the tile's own cells, never a zero or
//     staged cell, and the zero cells do not overlap the data cells.
// Prints "ok T=<T>" per tile edge and exits 0, or prints the first violation and
// exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../messi_amd/csrc/wave_layout.h"

enum Kind { GARBAGE = 0, ZERO, FINITE, CELL_S, CELL_F, CELL_L };

template <int T> static int check() {
  typedef messi::wave::Layout<T> X;
  const int N = X::END;
  std::vector<int> kind(N, GARBAGE), cell(N, -1);  // cell: finalisation step of a tile cell
  auto set = [&](int i, int k, int c = -1) {
    if (i < 0 || i >= N) {
      std::printf("T=%d: region cell %d out of [0, %d)\n", T, i, N);
      std::exit(1);
    }
    if (kind[i] != GARBAGE) {
      std::printf("T=%d: cell %d assigned twice (%d, %d)\n", T, i, kind[i], k);
      std::exit(1);
    }
    kind[i] = k;
    cell[i] = c;
  };
  auto own = [](int a, int b) { return b + T - 1 - a; };
  // staged part: Ld data (finite values from the table, zero cells included), its
  // pads and pad column; Sd and cP data (finite); their pad columns stay garbage
  for (int k = 0; k < X::N_ZERO_STAGED; ++k) set(X::zero_staged(k), ZERO);
  for (int a = 0; a < T; ++a)
    for (int c = 0; c < T; ++c) {
      set(X::ld(a, c), FINITE);
      set(X::sd(a, c), FINITE);
      set(X::cp(a, c), FINITE);
    }
  if (X::STAGED > N) return std::printf("T=%d: staged part beyond the end\n", T), 1;
  // wave part: zero cells, the bottom border row (finite), the tile's cells
  for (int k = 0; k < X::N_ZERO_WAVE; ++k) {
    const int i = X::zero_wave(k);
    if (i < X::STAGED) return std::printf("T=%d: wave zero cell %d in the staged part\n", T, i), 1;
    set(i, ZERO);
  }
  for (int b = 0; b < T; ++b) set(X::cs(T, b), FINITE);
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b) {
      set(X::cs(a, b), CELL_S, own(a, b));
      set(X::cf(a, b), CELL_F, own(a, b));
      set(X::cl(a, b), CELL_L, own(a, b));
    }
  // a cell holds a finite value at step st when it is a zero or finite cell, or
  // a tile cell finalised before st
  auto finite_at = [&](int i, int st) {
    const int k = kind[i];
    return k == ZERO || k == FINITE || ((k == CELL_S || k == CELL_F || k == CELL_L) && cell[i] < st);
  };
  auto in = [&](int i) { return i >= 0 && i < N; };
  long needed = 0, zero = 0, dead = 0;
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b)
      for (int st = 0; st <= X::LAST; ++st) {
        const int l1 = X::base_col_l(a, b) + X::off_col_l(st), s1 = X::base_col_s(a, b) + X::off_col_s(st);
        const int l2 = X::base_row_l(a, b) + X::off_row_l(st), s2 = X::base_row_s(a, b) + X::off_row_s(st);
        if (!in(l1) || !in(s1) || !in(l2) || !in(s2))
          return std::printf("T=%d: lane (%d,%d) step %d reads out of bounds\n", T, a, b, st), 1;
        const bool is_dead = st > own(a, b);
        // column term
        const int ap = b + T - st;
        if (is_dead) {
          ++dead;
        } else if (ap > a && ap <= T) {
          ++needed;
          const bool ok = l1 == X::ld(a, ap - 1) && s1 == X::cs(ap, b) && kind[l1] == FINITE &&
                          (ap == T ? kind[s1] == FINITE : (kind[s1] == CELL_S && cell[s1] == st - 1));
          if (!ok) return std::printf("T=%d: column term of (%d,%d) at step %d\n", T, a, b, st), 1;
        } else if (ap > T) {
          ++zero;
          if (!((kind[s1] == ZERO && finite_at(l1, st)) || (kind[l1] == ZERO && finite_at(s1, st))))
            return std::printf("T=%d: zero column term of (%d,%d) at step %d not zero\n", T, a, b, st), 1;
        } else {
          return std::printf("T=%d: unclassified column term (%d,%d) step %d\n", T, a, b, st), 1;
        }
        // row term
        const int bp = st - T + a;
        if (is_dead) {
          ++dead;
        } else if (bp >= 0 && bp < b) {
          ++needed;
          const bool ok = l2 == X::cl(a, bp) && s2 == X::sd(bp + 1, b) && kind[l2] == CELL_L &&
                          cell[l2] == st - 1 && kind[s2] == FINITE;
          if (!ok) return std::printf("T=%d: row term of (%d,%d) at step %d\n", T, a, b, st), 1;
        } else if (bp < 0) {
          ++zero;
          if (!((kind[l2] == ZERO && finite_at(s2, st)) || (kind[s2] == ZERO && finite_at(l2, st))))
            return std::printf("T=%d: zero row term of (%d,%d) at step %d not zero\n", T, a, b, st), 1;
        } else {
          return std::printf("T=%d: unclassified row term (%d,%d) step %d\n", T, a, b, st), 1;
        }
      }
  // every needed term is seen exactly once per cell: sum over cells of own(a,b) + 1
  // (the column terms a < ap <= T and the row terms 0 <= bp < b)
  long want = 0;
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b) want += (T - 1 - a) + 1 + b;
  if (needed != want) return std::printf("T=%d: %ld needed terms, expected %ld\n", T, needed, want), 1;
  std::printf("ok T=%d: %d cells, %ld needed, %ld zero, %ld dead reads\n", T, N, needed, zero, dead);
  return 0;
}

int main() {
  if (check<16>()) return 1;
  if (check<32>()) return 1;
  if (check<8>()) return 1;
  return 0;
}
