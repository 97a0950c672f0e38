/* Cross-checks of the CPU oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitize.py: random small inputs (duplicate and nested hyperedges, empty rows, equal
 * timestamps excluded), every restatement against its brute-force definition. Exit 0 = all agree
 * and no sanitizer report. Test infrastructure only. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_static_mochy(int32_t E, const int64_t* off, const int32_t* nodes, int64_t out[30], int threads);
int oracle_static_brute(int32_t E, const int64_t* off, const int32_t* nodes, int64_t out[30]);
int oracle_thyme_temporal(int32_t E, const int64_t* off, const int32_t* nodes, const int64_t* times,
                          int64_t delta, int64_t out[96]);
int oracle_temporal_brute(int32_t E, const int64_t* off, const int32_t* nodes, const int64_t* times,
                          int64_t delta, int64_t out[96]);
int oracle_strict_inc(const int64_t* const off[3], const int32_t* const nodes[3], const int32_t n[3],
                      int64_t out[30]);
int oracle_strict_inc_brute(const int64_t* const off[3], const int32_t* const nodes[3],
                            const int32_t n[3], int64_t out[30]);

static uint64_t rng_state = 88172645463325252ull;
static uint32_t rnd(uint32_t n) {  /* xorshift64 */
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (uint32_t)(rng_state % n);
}

/* E rows over V nodes, sizes 0..kmax (rows may repeat a node; the oracle de-duplicates), some
 * rows copied from earlier ones. */
static void random_rows(int32_t E, int32_t V, int kmax, int64_t* off, int32_t* nodes) {
  off[0] = 0;
  for (int32_t e = 0; e < E; ++e) {
    int64_t w = off[e];
    if (e > 0 && rnd(5) == 0) {
      int32_t src = (int32_t)rnd((uint32_t)e);
      for (int64_t q = off[src]; q < off[src + 1]; ++q) nodes[w++] = nodes[q];
    } else {
      int k = (int)rnd((uint32_t)kmax + 1);
      for (int q = 0; q < k; ++q) nodes[w++] = (int32_t)rnd((uint32_t)V);
    }
    off[e + 1] = w;
  }
}

static int same(const int64_t* a, const int64_t* b, int n, const char* what, int round) {
  if (memcmp(a, b, sizeof(int64_t) * (size_t)n) == 0) return 1;
  fprintf(stderr, "%s differs in round %d\n", what, round);
  return 0;
}

int main(void) {
  enum { EMAX = 48, KMAX = 7 };
  int64_t off[3][EMAX + 1], times[EMAX];
  int32_t nodes[3][EMAX * KMAX];
  int64_t a[96], b[96];
  for (int round = 0; round < 40; ++round) {
    int32_t E = 3 + (int32_t)rnd(EMAX - 2), V = 2 + (int32_t)rnd(20);
    random_rows(E, V, KMAX, off[0], nodes[0]);
    if (oracle_static_mochy(E, off[0], nodes[0], a, 2) || oracle_static_brute(E, off[0], nodes[0], b)) return 2;
    if (!same(a, b, 30, "static", round)) return 1;
    int64_t t = 0;
    for (int32_t e = 0; e < E; ++e) times[e] = (t += 1 + rnd(9));
    for (int32_t e = E - 1; e > 0; --e) {  /* input order is not time order */
      int32_t j = (int32_t)rnd((uint32_t)e + 1);
      int64_t x = times[e]; times[e] = times[j]; times[j] = x;
    }
    int64_t delta = 1 + rnd(60);
    if (oracle_thyme_temporal(E, off[0], nodes[0], times, delta, a) ||
        oracle_temporal_brute(E, off[0], nodes[0], times, delta, b)) return 2;
    if (!same(a, b, 96, "temporal", round)) return 1;
    int32_t n[3];
    const int64_t* po[3];
    const int32_t* pn[3];
    for (int l = 0; l < 3; ++l) {
      n[l] = 1 + (int32_t)rnd(EMAX);
      random_rows(n[l], V, KMAX, off[l], nodes[l]);
      po[l] = off[l];
      pn[l] = nodes[l];
    }
    if (oracle_strict_inc(po, pn, n, a) || oracle_strict_inc_brute(po, pn, n, b)) return 2;
    if (!same(a, b, 30, "strict-inc", round)) return 1;
  }
  printf("oracle sanitize: 40 rounds agree\n");
  return 0;
}
