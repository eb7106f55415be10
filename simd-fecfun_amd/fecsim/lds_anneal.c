/* Simulated annealing of the row-pair decoder's LDS layout (driven by lds_layout.py).

   A wave holds 64 / L codeword slots of L = Z/2 lanes; item (slot, lr) is a lane, item
   (slot, u) a pair position.  For a shift class c the lane of item (slot, lr) reads pair
   position (c + lr) mod L of its slot.  A layout puts each lane item in the first or
   second half-wave (32 lanes each; the one spare lane of the second half mirrors another
   lane of it, a broadcast) and each position item at one of 64 dwords (bank = dword mod
   32).  A half-wave access takes as many cycles as the most items requested from one
   bank; the cost is the sum over shift classes of (multiplicity x both halves' cycles).

   stdin:  L, then per class "c:multiplicity" pairs.
   args:   seed, steps.
   stdout: "cost <weighted cycles>", "pass" and the 0/1 half of each lane item,
           "dword" and the dword of each position item. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXI 64
static int L, NI, mult[32], cls[32], ncls;
static int in_first[MAXI], dw[MAXI];

static double urand(void) { return rand() / (RAND_MAX + 1.0); }

/* weighted cycles, and the sum over banks of pairs of items sharing one (a smoother
   guide for the search) */
static void score(long *hard, long *soft) {
  long h = 0, s = 0;
  for (int k = 0; k < ncls; ++k) {
    const int c = cls[k];
    for (int half = 0; half < 2; ++half) {
      int cnt[32] = {0};
      for (int it = 0; it < NI; ++it) {
        if (in_first[it] != (half == 0)) continue;
        const int sl = it / L, lr = it % L;
        cnt[dw[sl * L + (c + lr) % L] & 31]++;
      }
      int mx = 0;
      for (int b = 0; b < 32; ++b) {
        if (cnt[b] > mx) mx = cnt[b];
        s += (long)mult[c] * (cnt[b] * (cnt[b] - 1) / 2);
      }
      h += (long)mult[c] * mx;
    }
  }
  *hard = h;
  *soft = s;
}

int main(int argc, char **argv) {
  const int seed = argc > 1 ? atoi(argv[1]) : 1;
  const long steps = argc > 2 ? atol(argv[2]) : 1000000;
  srand(seed);
  if (scanf("%d", &L) != 1 || L < 2 || L > 32) return 1;
  NI = (64 / L) * L;
  int c, m;
  while (scanf(" %d:%d", &c, &m) == 2)
    if (c >= 0 && c < L) mult[c] = m;
  for (int i = 0; i < L; ++i)
    if (mult[i]) cls[ncls++] = i;
  const int first = NI < 32 ? NI : 32;   /* lane items in the first half-wave */
  /* random start */
  int perm[64];
  for (int i = 0; i < 64; ++i) perm[i] = i;
  for (int i = 63; i > 0; --i) {
    const int j = rand() % (i + 1), t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
  for (int i = 0; i < NI; ++i) dw[i] = perm[i];
  int nfree = 64 - NI, freed[64];
  for (int i = 0; i < nfree; ++i) freed[i] = perm[NI + i];
  int order[MAXI];
  for (int i = 0; i < NI; ++i) order[i] = i;
  for (int i = NI - 1; i > 0; --i) {
    const int j = rand() % (i + 1), t = order[i];
    order[i] = order[j];
    order[j] = t;
  }
  for (int i = 0; i < NI; ++i) in_first[order[i]] = i < first;
  long h, s;
  score(&h, &s);
  double cur = h + 0.05 * s, best = 1e300;
  long best_h = h;
  int best_first[MAXI], best_dw[MAXI];
  memcpy(best_first, in_first, sizeof in_first);
  memcpy(best_dw, dw, sizeof dw);
  const double t0 = 3.0, t1 = 0.02;
  double temp = t0;
  const double decay = pow(t1 / t0, 1.0 / (double)steps);
  for (long st = 0; st < steps; ++st, temp *= decay) {
    const int mv = rand() % 3;
    int a = 0, b = 0, old = 0, fi = 0;
    if (mv == 0) {   /* swap two lane items between the halves */
      if (first == NI) continue;
      do a = rand() % NI; while (!in_first[a]);
      do b = rand() % NI; while (in_first[b]);
      in_first[a] = 0;
      in_first[b] = 1;
    } else if (mv == 1) {   /* swap the dwords of two position items */
      a = rand() % NI;
      b = rand() % NI;
      if (a == b) continue;
      const int t = dw[a];
      dw[a] = dw[b];
      dw[b] = t;
    } else {   /* move a position item to a free dword */
      if (!nfree) continue;
      a = rand() % NI;
      fi = rand() % nfree;
      old = dw[a];
      dw[a] = freed[fi];
      freed[fi] = old;
    }
    score(&h, &s);
    const double nw = h + 0.05 * s;
    if (nw <= cur || urand() < exp((cur - nw) / temp)) {
      cur = nw;
      if (nw < best) {
        best = nw;
        best_h = h;
        memcpy(best_first, in_first, sizeof in_first);
        memcpy(best_dw, dw, sizeof dw);
      }
    } else if (mv == 0) {
      in_first[a] = 1;
      in_first[b] = 0;
    } else if (mv == 1) {
      const int t = dw[a];
      dw[a] = dw[b];
      dw[b] = t;
    } else {
      freed[fi] = dw[a];
      dw[a] = old;
    }
  }
  printf("cost %ld\npass", best_h);
  for (int i = 0; i < NI; ++i) printf(" %d", best_first[i] ? 0 : 1);
  printf("\ndword");
  for (int i = 0; i < NI; ++i) printf(" %d", best_dw[i]);
  printf("\n");
  return 0;
}
