/*
 * hm_oracle.c — CPU ORACLE for exact hypergraph-motif counting. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code, and
 * only as the checker (or as the timed CPU baseline). The product path (libhmotif) never calls it.
 *
 * What it restates, and what pins each function:
 *   oracle_static_mochy    MoCHy-E, MoCHy-master/main_exact_par.cpp:74-173 (projection by stamp
 *                          array :74-89, sorted rows :91-101, wedge loop :122-165), classification
 *                          motif_id.cpp:24-34 + id_to_index :5-22, reader rule read_data.cpp:34.
 *                          Pinned by: golden vectors produced by the reference MoCHy binaries
 *                          (tests/golden), oracle_static_brute, closed forms (stars, paths,
 *                          identical edges) and the wedge invariant
 *                          sum(open) + 3 sum(closed) = sum_a C(d_a, 2).
 *   oracle_static_brute    definition: every unordered triple of distinct hyperedges with at least
 *                          two overlapping pairs, classified by the emptiness of its 7 Venn regions.
 *                          Pinned by the same goldens on small inputs and by the closed forms.
 *   oracle_thyme_temporal  THyMe+, THyMe-main/code/main_thymeP.cpp:54-389 (update_add :54-89,
 *                          update_delete :91-115, static dedupe :164-179, anchor loop :219-367),
 *                          motif_id.cpp:25-42,57-108. Defined only for distinct timestamps (with
 *                          ties THyMe is nondeterministic: read_data.cpp:35-36, motif_id.cpp:68-78);
 *                          here ties are broken by input order. Pinned by: golden vectors from the
 *                          reference THyMe+ binary on distinct-timestamp inputs, oracle_temporal_brute,
 *                          and closed forms (m copies of one edge inside delta -> C(m,3) in motif 96).
 *                          An empty hyperedge is in no instance (THyMe's reader cannot read one).
 *   oracle_thyme_temporal_range  the same count restricted to the instances whose newest edge has
 *                          a time-order position in [lo, hi): shards of one count for a process
 *                          pool. Pinned by: the shards of every temporal golden summing to the
 *                          whole count (tests/test_oracle.py).
 *   oracle_temporal_brute  definition: every triple of temporal hyperedges whose time span is
 *                          <= delta and with at least two overlapping pairs (copies of one node set
 *                          overlap), classified by the Venn pattern written in time order.
 *   oracle_projection_degrees  d_a = number of other hyperedges sharing a node with a.
 *   oracle_strict_inc      the reference's three-snapshot strict-increasing count,
 *                          temporalDynamicHMotif/src/temporal_count.cu:91-171 (candidates i < j from
 *                          the older snapshot's node lists of j's nodes :117-131, k > j from the
 *                          newest's :133-147, pair checks and classification :149-170 with
 *                          kernel/motif_utils.cuh:80-100), without the kernel's 1,024-candidate cap
 *                          (:79-80). Pinned by: the reference driver's synthetic window
 *                          (temporalDynamicHMotif/src/main.cu:267-305: bin 26 = 1, then delta
 *                          bin 10 +1 / bin 26 -1) and oracle_strict_inc_brute.
 *   oracle_strict_inc_brute  definition: every (i, j, k), i < j < k, older row i, middle row j,
 *                          newest row k, all three pairs overlapping, classified by its Venn pattern.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* MoCHy numbering (MoCHy-master/motif_id.cpp:5-22): 7-bit Venn-emptiness id -> motif index 1..30
 * (0 = not a connected triple). */
static const int MOCHY_ID_TO_INDEX[128] = {
    0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
    0,  0,  21, 23, 22, 24, 23, 25, 24, 26, 0,  0,  0,  0,  0,  0,  0,  0,  21, 22, 23, 24,
    23, 24, 25, 26, 21, 23, 23, 25, 22, 24, 24, 26, 27, 28, 28, 29, 28, 29, 29, 30, 1,  2,
    2,  3,  2,  3,  3,  4,  5,  6,  6,  8,  7,  9,  9,  10, 5,  7,  6,  9,  6,  9,  8,  10,
    11, 13, 12, 14, 13, 15, 14, 16, 5,  6,  7,  9,  6,  8,  9,  10, 11, 12, 13, 14, 13, 14,
    15, 16, 11, 13, 13, 15, 12, 14, 14, 16, 17, 18, 18, 19, 18, 19, 19, 20};

/* THyMe temporal numbering (THyMe-main/code/motif_id.cpp:25-42): time-ordered id -> 1..96. */
static const int THYME_ID_TO_INDEX_TEMPORAL[128] = {
    0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
    0,  0,  55, 61, 58, 67, 62, 73, 68, 76, 0,  0,  0,  0,  0,  0,  0,  0,  56, 59, 63, 69,
    64, 70, 74, 77, 57, 65, 66, 75, 60, 71, 72, 78, 79, 80, 81, 83, 82, 84, 85, 86, 96, 87,
    88, 1,  89, 2,  3,  4,  90, 5,  6,  11, 93, 14, 15, 20, 91, 94, 7,  16, 8,  17, 12, 21,
    23, 29, 26, 35, 30, 41, 36, 44, 92, 9,  95, 18, 10, 13, 19, 22, 24, 27, 31, 37, 32, 38,
    42, 45, 25, 33, 34, 43, 28, 39, 40, 46, 47, 48, 49, 51, 50, 52, 53, 54};

const int* oracle_mochy_table(void) { return MOCHY_ID_TO_INDEX; }
const int* oracle_thyme_temporal_table(void) { return THYME_ID_TO_INDEX_TEMPORAL; }

/* Venn-region emptiness id of (x1, x2, x3) from sizes, pairwise overlaps and the triple overlap
 * (MoCHy-master/motif_id.cpp:24-34). Pair bits: x1x2 -> 8, x2x3 -> 16, x3x1 -> 32. */
static int venn_id(int s1, int s2, int s3, int c12, int c23, int c31, int g) {
  int r1 = s1 - (c12 + c31) + g;
  int r2 = s2 - (c23 + c12) + g;
  int r3 = s3 - (c31 + c23) + g;
  int r12 = c12 - g, r23 = c23 - g, r31 = c31 - g;
  return (r1 > 0) + ((r2 > 0) << 1) + ((r3 > 0) << 2) + ((r12 > 0) << 3) + ((r23 > 0) << 4) +
         ((r31 > 0) << 5) + ((g > 0) << 6);
}

static int mochy_bin(int sa, int sb, int sc, int cab, int cbc, int cca, int g) {
  return MOCHY_ID_TO_INDEX[venn_id(sa, sb, sc, cab, cbc, cca, g)] - 1;
}

/* ---------------------------------------------------------------------------------------------
 * Shared helpers: per-hyperedge node lists with repeats removed (MoCHy-master/read_data.cpp:34),
 * sorted for intersections; node -> hyperedge lists in increasing hyperedge order.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
  int32_t E, V;
  int64_t* off;    /* E + 1 */
  int32_t* nodes;  /* sorted, unique per edge */
  int64_t* voff;   /* V + 1 */
  int32_t* vedges; /* increasing */
} csr_t;

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

static void csr_free(csr_t* g) {
  free(g->off);
  free(g->nodes);
  free(g->voff);
  free(g->vedges);
}

static int csr_build(int32_t E, const int64_t* off, const int32_t* nodes, csr_t* g) {
  memset(g, 0, sizeof(*g));
  g->E = E;
  int64_t I = off[E];
  int32_t vmax = -1;
  for (int64_t i = 0; i < I; ++i)
    if (nodes[i] > vmax) vmax = nodes[i];
  g->V = vmax + 1;
  g->off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(E + 1));
  g->nodes = (int32_t*)malloc(sizeof(int32_t) * (size_t)(I > 0 ? I : 1));
  g->voff = (int64_t*)calloc((size_t)g->V + 1, sizeof(int64_t));
  if (!g->off || !g->nodes || !g->voff) return -1;
  int64_t w = 0;
  g->off[0] = 0;
  for (int32_t e = 0; e < E; ++e) {
    int64_t s = w;
    for (int64_t i = off[e]; i < off[e + 1]; ++i) g->nodes[w++] = nodes[i];
    qsort(g->nodes + s, (size_t)(w - s), sizeof(int32_t), cmp_i32);
    int64_t u = s;
    for (int64_t i = s; i < w; ++i)
      if (i == s || g->nodes[i] != g->nodes[i - 1]) g->nodes[u++] = g->nodes[i];
    w = u;
    g->off[e + 1] = w;
  }
  for (int64_t i = 0; i < w; ++i) g->voff[g->nodes[i] + 1]++;
  for (int32_t v = 0; v < g->V; ++v) g->voff[v + 1] += g->voff[v];
  g->vedges = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w > 0 ? w : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(g->V > 0 ? g->V : 1));
  if (!g->vedges || !fill) return -1;
  for (int32_t v = 0; v < g->V; ++v) fill[v] = g->voff[v];
  for (int32_t e = 0; e < E; ++e)
    for (int64_t i = g->off[e]; i < g->off[e + 1]; ++i) g->vedges[fill[g->nodes[i]]++] = e;
  free(fill);
  return 0;
}

static int inter_size(const int32_t* a, int na, const int32_t* b, int nb) {
  int i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else { ++c; ++i; ++j; }
  }
  return c;
}

/* Projection rows: for every hyperedge a, sorted (b, C_ab) for all b != a sharing a node.
 * Follows the stamp-array construction of MoCHy-master/main_exact_par.cpp:74-101. */
typedef struct {
  int64_t* roff;   /* E + 1 */
  int32_t* rid;    /* neighbour id, sorted per row */
  int32_t* rcnt;   /* C_ab */
} proj_t;

typedef struct { int32_t id, cnt; } pair_t;
static int cmp_pair(const void* a, const void* b) {
  int32_t x = ((const pair_t*)a)->id, y = ((const pair_t*)b)->id;
  return (x > y) - (x < y);
}

static int proj_build(const csr_t* g, proj_t* p) {
  int32_t E = g->E;
  p->roff = (int64_t*)calloc((size_t)E + 1, sizeof(int64_t));
  int64_t* stamp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(E > 0 ? E : 1));
  if (!p->roff || !stamp) return -1;
  /* pass 1: degrees */
  for (int32_t e = 0; e < E; ++e) stamp[e] = -1;
  for (int32_t a = 0; a < E; ++a) {
    int64_t d = 0;
    for (int64_t i = g->off[a]; i < g->off[a + 1]; ++i) {
      int32_t v = g->nodes[i];
      for (int64_t k = g->voff[v]; k < g->voff[v + 1]; ++k) {
        int32_t b = g->vedges[k];
        if (b == a || stamp[b] == a) continue;
        stamp[b] = a;
        ++d;
      }
    }
    p->roff[a + 1] = p->roff[a] + d;
  }
  int64_t P = p->roff[E];
  p->rid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(P > 0 ? P : 1));
  p->rcnt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(P > 0 ? P : 1));
  pair_t* tmp = (pair_t*)malloc(sizeof(pair_t) * (size_t)(E > 0 ? E : 1));
  int64_t* slot = (int64_t*)malloc(sizeof(int64_t) * (size_t)(E > 0 ? E : 1));
  if (!p->rid || !p->rcnt || !tmp || !slot) return -1;
  /* pass 2: (a << 31) + slot stamps as in main_exact_par.cpp:82-86 */
  for (int32_t e = 0; e < E; ++e) stamp[e] = -1;
  for (int32_t a = 0; a < E; ++a) {
    int32_t d = 0;
    for (int64_t i = g->off[a]; i < g->off[a + 1]; ++i) {
      int32_t v = g->nodes[i];
      for (int64_t k = g->voff[v]; k < g->voff[v + 1]; ++k) {
        int32_t b = g->vedges[k];
        if (b == a) continue;
        if (stamp[b] != a) {
          stamp[b] = a;
          slot[b] = d;
          tmp[d].id = b;
          tmp[d].cnt = 0;
          ++d;
        }
        tmp[slot[b]].cnt++;
      }
    }
    qsort(tmp, (size_t)d, sizeof(pair_t), cmp_pair);
    for (int32_t k = 0; k < d; ++k) {
      p->rid[p->roff[a] + k] = tmp[k].id;
      p->rcnt[p->roff[a] + k] = tmp[k].cnt;
    }
  }
  free(stamp);
  free(tmp);
  free(slot);
  return 0;
}

static void proj_free(proj_t* p) {
  free(p->roff);
  free(p->rid);
  free(p->rcnt);
}

/* C_bc by binary search in the sorted row of b (MoCHy looks it up in a hash map,
 * main_exact_par.cpp:143-146; the value is the same). */
static int32_t proj_lookup(const proj_t* p, int32_t b, int32_t c) {
  int64_t lo = p->roff[b], hi = p->roff[b + 1];
  while (lo < hi) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (p->rid[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < p->roff[b + 1] && p->rid[lo] == c) ? p->rcnt[lo] : 0;
}

int oracle_projection_degrees(int32_t E, const int64_t* off, const int32_t* nodes, int64_t* deg) {
  csr_t g;
  proj_t p;
  if (csr_build(E, off, nodes, &g) || proj_build(&g, &p)) return -1;
  for (int32_t a = 0; a < E; ++a) deg[a] = p.roff[a + 1] - p.roff[a];
  proj_free(&p);
  csr_free(&g);
  return 0;
}

/* MoCHy-E: MoCHy-master/main_exact_par.cpp:122-165. For every centre a and every pair i < j of
 * its projected neighbours (b, c): if C_bc > 0 the triple is closed and is counted only when
 * a < b (rows are sorted, so a is the minimum); otherwise it is open and counted at its unique
 * centre a. g_abc = |(a ∩ b) ∩ c|. */
int oracle_static_mochy(int32_t E, const int64_t* off, const int32_t* nodes, int64_t out[30],
                        int nthreads) {
  csr_t g;
  proj_t p;
  for (int i = 0; i < 30; ++i) out[i] = 0;
  if (E <= 0) return 0;
  if (csr_build(E, off, nodes, &g) || proj_build(&g, &p)) return -1;
  if (nthreads <= 0) nthreads = 1;
  int64_t total[30] = {0};
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    int64_t h[30] = {0};
    int32_t* inter = (int32_t*)malloc(sizeof(int32_t) * 1024);
    int icap = 1024;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int32_t a = 0; a < E; ++a) {
      const int32_t* na = g.nodes + g.off[a];
      int sa = (int)(g.off[a + 1] - g.off[a]);
      if (sa > icap) {
        icap = sa;
        inter = (int32_t*)realloc(inter, sizeof(int32_t) * (size_t)icap);
      }
      for (int64_t i = p.roff[a]; i < p.roff[a + 1]; ++i) {
        int32_t b = p.rid[i];
        int cab = p.rcnt[i];
        const int32_t* nb = g.nodes + g.off[b];
        int sb = (int)(g.off[b + 1] - g.off[b]);
        /* a ∩ b (main_exact_par.cpp:135-136) */
        int cnt = 0, x = 0, y = 0;
        while (x < sa && y < sb) {
          if (na[x] < nb[y]) ++x;
          else if (na[x] > nb[y]) ++y;
          else { inter[cnt++] = na[x]; ++x; ++y; }
        }
        for (int64_t j = i + 1; j < p.roff[a + 1]; ++j) {
          int32_t c = p.rid[j];
          int cca = p.rcnt[j];
          const int32_t* nc = g.nodes + g.off[c];
          int sc = (int)(g.off[c + 1] - g.off[c]);
          int cbc = proj_lookup(&p, b, c);
          if (cbc) {
            if (a < b) {
              int gabc = inter_size(inter, cnt, nc, sc);
              h[mochy_bin(sa, sb, sc, cab, cbc, cca, gabc)]++;
            }
          } else {
            h[mochy_bin(sa, sb, sc, cab, 0, cca, 0)]++;
          }
        }
      }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (int k = 0; k < 30; ++k) total[k] += h[k];
    free(inter);
  }
  for (int k = 0; k < 30; ++k) out[k] = total[k];
  proj_free(&p);
  csr_free(&g);
  return 0;
}

/* Brute force over all unordered triples (definition of an h-motif instance). */
int oracle_static_brute(int32_t E, const int64_t* off, const int32_t* nodes, int64_t out[30]) {
  csr_t g;
  for (int i = 0; i < 30; ++i) out[i] = 0;
  if (E < 3) return 0;
  if (csr_build(E, off, nodes, &g)) return -1;
  int32_t* ab = (int32_t*)malloc(sizeof(int32_t) * 4096);
  for (int32_t a = 0; a < E; ++a) {
    const int32_t* na = g.nodes + g.off[a];
    int sa = (int)(g.off[a + 1] - g.off[a]);
    for (int32_t b = a + 1; b < E; ++b) {
      const int32_t* nb = g.nodes + g.off[b];
      int sb = (int)(g.off[b + 1] - g.off[b]);
      int nab = 0, x = 0, y = 0;
      while (x < sa && y < sb) {
        if (na[x] < nb[y]) ++x;
        else if (na[x] > nb[y]) ++y;
        else { ab[nab++] = na[x]; ++x; ++y; }
      }
      for (int32_t c = b + 1; c < E; ++c) {
        const int32_t* nc = g.nodes + g.off[c];
        int sc = (int)(g.off[c + 1] - g.off[c]);
        int cbc = inter_size(nb, sb, nc, sc);
        int cca = inter_size(nc, sc, na, sa);
        int links = (nab > 0) + (cbc > 0) + (cca > 0);
        if (links < 2) continue;
        int gabc = inter_size(ab, nab, nc, sc);
        out[mochy_bin(sa, sb, sc, nab, cbc, cca, gabc)]++;
      }
    }
  }
  free(ab);
  csr_free(&g);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Temporal (THyMe+).
 * ------------------------------------------------------------------------------------------- */
static int temporal_bin(int s1, int s2, int s3, int c12, int c23, int c31, int g) {
  return THYME_ID_TO_INDEX_TEMPORAL[venn_id(s1, s2, s3, c12, c23, c31, g)] - 1;
}

typedef struct {
  const int64_t* t;
  int32_t idx;
} tkey_t;

static const csr_t* g_sort_csr; /* comparator context (single-threaded use) */
static const int64_t* g_sort_times;

static int cmp_time(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  int64_t tx = g_sort_times[x], ty = g_sort_times[y];
  if (tx != ty) return (tx > ty) - (tx < ty);
  return (x > y) - (x < y);
}

static int cmp_nodeset(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  const csr_t* g = g_sort_csr;
  int64_t nx = g->off[x + 1] - g->off[x], ny = g->off[y + 1] - g->off[y];
  const int32_t* px = g->nodes + g->off[x];
  const int32_t* py = g->nodes + g->off[y];
  int64_t n = nx < ny ? nx : ny;
  for (int64_t i = 0; i < n; ++i)
    if (px[i] != py[i]) return (px[i] > py[i]) - (px[i] < py[i]);
  if (nx != ny) return (nx > ny) - (nx < ny);
  return (x > y) - (x < y);
}

/* growable int arrays */
typedef struct { int32_t* v; int32_t n, cap; } ivec;
static void iv_push(ivec* a, int32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? a->cap * 2 : 8;
    a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}
static void iv_remove_value(ivec* a, int32_t x) {
  for (int32_t i = 0; i < a->n; ++i)
    if (a->v[i] == x) {
      memmove(a->v + i, a->v + i + 1, sizeof(int32_t) * (size_t)(a->n - i - 1));
      a->n--;
      return;
    }
}

/* number of pairs (x in X, y in Y) with x < y, X and Y ascending (main_thymeP.cpp:315-322) */
static int64_t ordered_pairs(const int64_t* X, int nx, const int64_t* Y, int ny) {
  int64_t c = 0;
  int j = 0;
  for (int i = 0; i < nx; ++i) {
    while (j < ny && !(X[i] < Y[j])) ++j;
    c += ny - j;
  }
  return c;
}

/* THyMe+ (THyMe-main/code/main_thymeP.cpp:54-389) over int64 timestamps, counting the instances
 * whose newest edge sits at a time-order position in [lo, hi). THyMe+ builds the window state edge
 * by edge; the state at position lo depends only on the edges inside lo's window, so the loop starts
 * at the window's first position (adding, never counting, until lo). The counts do not depend on
 * the order of the adjacency lists (closed pairs are taken once by static id, open pairs add both
 * time orders), so any split of [0, E) into ranges sums to the whole count. */
static int thyme_range(int32_t E, const int64_t* off, const int32_t* nodes, const int64_t* times,
                       int64_t delta, int32_t lo, int32_t hi, int64_t out[96]) {
  for (int i = 0; i < 96; ++i) out[i] = 0;
  if (E <= 0 || lo >= hi) return 0;
  csr_t g;
  if (csr_build(E, off, nodes, &g)) return -1;
  /* static dedupe by node set (main_thymeP.cpp:164-179) */
  int32_t* byset = (int32_t*)malloc(sizeof(int32_t) * (size_t)E);
  int32_t* sid = (int32_t*)malloc(sizeof(int32_t) * (size_t)E);
  for (int32_t i = 0; i < E; ++i) byset[i] = i;
  g_sort_csr = &g;
  qsort(byset, (size_t)E, sizeof(int32_t), cmp_nodeset);
  int32_t S = 0;
  int32_t* rep = (int32_t*)malloc(sizeof(int32_t) * (size_t)E); /* static id -> a temporal edge */
  for (int32_t k = 0; k < E; ++k) {
    int32_t e = byset[k];
    if (k > 0) {
      int32_t p = byset[k - 1];
      int64_t np = g.off[p + 1] - g.off[p], ne = g.off[e + 1] - g.off[e];
      int same = (np == ne) &&
                 memcmp(g.nodes + g.off[p], g.nodes + g.off[e], sizeof(int32_t) * (size_t)ne) == 0;
      if (!same) ++S;
    }
    sid[e] = S;
    rep[S] = e;
  }
  ++S;
  /* time order (main_thymeP.cpp:190-195), ties broken by input index */
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)E);
  for (int32_t i = 0; i < E; ++i) order[i] = i;
  g_sort_times = times;
  qsort(order, (size_t)E, sizeof(int32_t), cmp_time);

  int32_t V = g.V;
  ivec* V2E = (ivec*)calloc((size_t)(V > 0 ? V : 1), sizeof(ivec));
  ivec* adj = (ivec*)calloc((size_t)S, sizeof(ivec));    /* neighbour static ids */
  ivec* adjc = (ivec*)calloc((size_t)S, sizeof(ivec));   /* their overlaps */
  /* E2times: times of the copies of each static edge inside the window, ascending */
  int64_t** etimes = (int64_t**)calloc((size_t)S, sizeof(int64_t*));
  int32_t* ehead = (int32_t*)calloc((size_t)S, sizeof(int32_t));
  int32_t* ecnt = (int32_t*)calloc((size_t)S, sizeof(int32_t));
  int32_t* ecap = (int32_t*)calloc((size_t)S, sizeof(int32_t));
  int64_t* stamp = (int64_t*)malloc(sizeof(int64_t) * (size_t)S);
  int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
  int64_t* markA = (int64_t*)malloc(sizeof(int64_t) * (size_t)S);  /* anchor stamp */
  int32_t* markC = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);  /* C_ac */
  int64_t* markB = (int64_t*)malloc(sizeof(int64_t) * (size_t)S);  /* c in N(b) stamp */
  for (int32_t s = 0; s < S; ++s) { stamp[s] = -1; markA[s] = -1; markB[s] = -1; }
  int32_t* inter = (int32_t*)malloc(sizeof(int32_t) * 4096);
  int64_t bstamp = 0;

#define SNODES(s) (g.nodes + g.off[rep[s]])
#define SSIZE(s) ((int)(g.off[rep[s] + 1] - g.off[rep[s]]))

  int32_t start = 0;
  while (times[order[start]] + delta < times[order[lo]]) ++start;
  for (int32_t end = start; end < hi; ++end) {
    int32_t ea = order[end];
    int32_t a = sid[ea];
    int64_t ta = times[ea];
    int sa = SSIZE(a);
    /* (1) evict (main_thymeP.cpp:228-234; update_delete :91-115) */
    while (times[order[start]] + delta < ta) {
      int32_t s = sid[order[start]];
      ehead[s]++;
      if (ehead[s] == ecnt[s]) {
        ehead[s] = ecnt[s] = 0;
        const int32_t* ns = SNODES(s);
        for (int k = 0; k < SSIZE(s); ++k) iv_remove_value(&V2E[ns[k]], s);
        for (int32_t k = 0; k < adj[s].n; ++k) {
          int32_t o = adj[s].v[k];
          for (int32_t q = 0; q < adj[o].n; ++q)
            if (adj[o].v[q] == s) {
              memmove(adj[o].v + q, adj[o].v + q + 1, sizeof(int32_t) * (size_t)(adj[o].n - q - 1));
              memmove(adjc[o].v + q, adjc[o].v + q + 1, sizeof(int32_t) * (size_t)(adjc[o].n - q - 1));
              adj[o].n--;
              adjc[o].n--;
              break;
            }
        }
        adj[s].n = 0;
        adjc[s].n = 0;
      }
      ++start;
    }
    /* (2) add (update_add :54-89) */
    if (ecnt[a] == ecap[a]) {
      ecap[a] = ecap[a] ? ecap[a] * 2 : 4;
      etimes[a] = (int64_t*)realloc(etimes[a], sizeof(int64_t) * (size_t)ecap[a]);
    }
    etimes[a][ecnt[a]++] = ta;
    if (ecnt[a] - ehead[a] == 1) {
      const int32_t* na = SNODES(a);
      for (int k = 0; k < sa; ++k) iv_push(&V2E[na[k]], a);
      int32_t d = 0;
      for (int k = 0; k < sa; ++k) {
        ivec* lst = &V2E[na[k]];
        for (int32_t q = 0; q < lst->n; ++q) {
          int32_t o = lst->v[q];
          if (o == a) continue;
          if (stamp[o] != end) {
            stamp[o] = end;
            slot[o] = d++;
            iv_push(&adj[a], o);
            iv_push(&adjc[a], 0);
          }
          adjc[a].v[slot[o]]++;
        }
      }
      for (int32_t k = 0; k < adj[a].n; ++k) {
        iv_push(&adj[adj[a].v[k]], a);
        iv_push(&adjc[adj[a].v[k]], adjc[a].v[k]);
      }
    }
    if (end < lo) continue; /* window state only */
    /* (3) instances with a as the newest edge (main_thymeP.cpp:240-335) */
    const int32_t* na = SNODES(a);
    int32_t da = adj[a].n;
    for (int32_t k = 0; k < da; ++k) {
      markA[adj[a].v[k]] = end;
      markC[adj[a].v[k]] = adjc[a].v[k];
    }
    const int64_t* Ta = etimes[a] + ehead[a];
    int na_t = ecnt[a] - ehead[a];
    for (int32_t i = 0; i < da; ++i) {
      int32_t b = adj[a].v[i];
      int cab = adjc[a].v[i];
      int sb = SSIZE(b);
      const int32_t* nb = SNODES(b);
      const int64_t* Tb = etimes[b] + ehead[b];
      int nb_t = ecnt[b] - ehead[b];
      int cnt = 0, x = 0, y = 0;
      while (x < sa && y < sb) {
        if (na[x] < nb[y]) ++x;
        else if (na[x] > nb[y]) ++y;
        else { inter[cnt++] = na[x]; ++x; ++y; }
      }
      ++bstamp;
      for (int32_t j = 0; j < adj[b].n; ++j) {
        int32_t c = adj[b].v[j];
        if (c == a) continue;
        markB[c] = bstamp;
        int cbc = adjc[b].v[j];
        int sc = SSIZE(c);
        int cca = (markA[c] == end) ? markC[c] : 0;
        if (cca && !(b < c)) continue;
        int gabc = inter_size(inter, cnt, SNODES(c), sc);
        const int64_t* Tc = etimes[c] + ehead[c];
        int nc_t = ecnt[c] - ehead[c];
        int64_t bc = ordered_pairs(Tb, nb_t, Tc, nc_t);
        int64_t cb = (int64_t)nb_t * nc_t - bc;
        /* get_motif_index(..., 3, 1, 2): order b < c < a; (..., 3, 2, 1): c < b < a */
        out[temporal_bin(sb, sc, sa, cbc, cca, cab, gabc)] += bc;
        out[temporal_bin(sc, sb, sa, cbc, cab, cca, gabc)] += cb;
      }
      for (int32_t j = i + 1; j < da; ++j) {
        int32_t c = adj[a].v[j];
        if (markB[c] == bstamp) continue;
        int cca = adjc[a].v[j];
        int sc = SSIZE(c);
        const int64_t* Tc = etimes[c] + ehead[c];
        int nc_t = ecnt[c] - ehead[c];
        int64_t bc = ordered_pairs(Tb, nb_t, Tc, nc_t);
        int64_t cb = (int64_t)nb_t * nc_t - bc;
        out[temporal_bin(sb, sc, sa, 0, cca, cab, 0)] += bc;
        out[temporal_bin(sc, sb, sa, 0, cab, cca, 0)] += cb;
      }
      /* motifs 87-95 (main_thymeP.cpp:338-363): two copies of a and one of b, or the reverse */
      int64_t ab = ordered_pairs(Ta, na_t - 1, Tb, nb_t);
      int64_t ba = (int64_t)(na_t - 1) * nb_t - ab;
      int64_t bb = (int64_t)nb_t * (nb_t - 1) / 2;
      out[temporal_bin(sa, sb, sa, cab, cab, sa, cab)] += ab;
      out[temporal_bin(sb, sa, sa, cab, sa, cab, cab)] += ba;
      out[temporal_bin(sb, sb, sa, sb, cab, cab, cab)] += bb;
    }
    /* motif 96 (main_thymeP.cpp:366-367); copies of an empty hyperedge share no node, so they
       form no instance (THyMe's reader cannot read an empty hyperedge; here it is defined) */
    if (sa > 0 && na_t >= 2) out[95] += (int64_t)(na_t - 1) * (na_t - 2) / 2;
  }
#undef SNODES
#undef SSIZE
  for (int32_t v = 0; v < V; ++v) free(V2E[v].v);
  for (int32_t s = 0; s < S; ++s) { free(adj[s].v); free(adjc[s].v); free(etimes[s]); }
  free(V2E); free(adj); free(adjc); free(etimes); free(ehead); free(ecnt); free(ecap);
  free(stamp); free(slot); free(markA); free(markC); free(markB); free(inter);
  free(byset); free(sid); free(rep); free(order);
  csr_free(&g);
  return 0;
}

int oracle_thyme_temporal(int32_t E, const int64_t* off, const int32_t* nodes,
                          const int64_t* times, int64_t delta, int64_t out[96]) {
  return thyme_range(E, off, nodes, times, delta, 0, E, out);
}

/* The instances whose newest edge has time-order position in [lo, hi) (ties by input index). */
int oracle_thyme_temporal_range(int32_t E, const int64_t* off, const int32_t* nodes,
                                const int64_t* times, int64_t delta, int32_t lo, int32_t hi,
                                int64_t out[96]) {
  if (lo < 0 || hi > E || lo > hi) return -1;
  return thyme_range(E, off, nodes, times, delta, lo, hi, out);
}

/* Brute force over all triples of temporal hyperedges (definition). */
int oracle_temporal_brute(int32_t E, const int64_t* off, const int32_t* nodes,
                          const int64_t* times, int64_t delta, int64_t out[96]) {
  for (int i = 0; i < 96; ++i) out[i] = 0;
  if (E < 3) return 0;
  csr_t g;
  if (csr_build(E, off, nodes, &g)) return -1;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)E);
  for (int32_t i = 0; i < E; ++i) order[i] = i;
  g_sort_times = times;
  qsort(order, (size_t)E, sizeof(int32_t), cmp_time);
  int32_t* ab = (int32_t*)malloc(sizeof(int32_t) * 4096);
  for (int32_t p = 0; p < E; ++p) {
    int32_t x = order[p];
    const int32_t* nx = g.nodes + g.off[x];
    int sx = (int)(g.off[x + 1] - g.off[x]);
    for (int32_t q = p + 1; q < E; ++q) {
      int32_t y = order[q];
      if (times[y] - times[x] > delta) break;
      const int32_t* ny = g.nodes + g.off[y];
      int sy = (int)(g.off[y + 1] - g.off[y]);
      int nxy = 0, i = 0, j = 0;
      while (i < sx && j < sy) {
        if (nx[i] < ny[j]) ++i;
        else if (nx[i] > ny[j]) ++j;
        else { ab[nxy++] = nx[i]; ++i; ++j; }
      }
      for (int32_t r = q + 1; r < E; ++r) {
        int32_t z = order[r];
        if (times[z] - times[x] > delta) break;
        const int32_t* nz = g.nodes + g.off[z];
        int sz = (int)(g.off[z + 1] - g.off[z]);
        int cyz = inter_size(ny, sy, nz, sz);
        int czx = inter_size(nz, sz, nx, sx);
        if ((nxy > 0) + (cyz > 0) + (czx > 0) < 2) continue;
        int gxyz = inter_size(ab, nxy, nz, sz);
        out[temporal_bin(sx, sy, sz, nxy, cyz, czx, gxyz)]++;
      }
    }
  }
  free(ab);
  free(order);
  csr_free(&g);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Three-snapshot strict-increasing count (the reference's temporal-window kernel).
 * ------------------------------------------------------------------------------------------- */
static int strict_inc_build(const int64_t* const off[3], const int32_t* const nodes[3],
                            const int32_t n[3], csr_t g[3]) {
  for (int l = 0; l < 3; ++l)
    if (csr_build(n[l], off[l], nodes[l], &g[l])) return -1;
  return 0;
}

int oracle_strict_inc(const int64_t* const off[3], const int32_t* const nodes[3], const int32_t n[3],
                      int64_t out[30]) {
  csr_t g[3];
  for (int i = 0; i < 30; ++i) out[i] = 0;
  if (strict_inc_build(off, nodes, n, g)) return -1;
  const csr_t *O = &g[0], *M = &g[1], *N = &g[2];
  int32_t* candI = (int32_t*)malloc(sizeof(int32_t) * (size_t)(O->E > 0 ? O->E : 1));
  int32_t* candK = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N->E > 0 ? N->E : 1));
  uint8_t* seenI = (uint8_t*)calloc((size_t)(O->E > 0 ? O->E : 1), 1);
  uint8_t* seenK = (uint8_t*)calloc((size_t)(N->E > 0 ? N->E : 1), 1);
  int32_t* jk = (int32_t*)malloc(sizeof(int32_t) * 4096);
  for (int32_t j = 0; j < M->E; ++j) {
    const int32_t* nj = M->nodes + M->off[j];
    int sj = (int)(M->off[j + 1] - M->off[j]);
    int ni = 0, nk = 0;
    for (int p = 0; p < sj; ++p) {  /* older rows through j's nodes, i < j */
      int32_t v = nj[p];
      if (v < O->V)
        for (int64_t q = O->voff[v]; q < O->voff[v + 1]; ++q) {
          int32_t i = O->vedges[q];
          if (i < j && !seenI[i]) { seenI[i] = 1; candI[ni++] = i; }
        }
      if (v < N->V)
        for (int64_t q = N->voff[v]; q < N->voff[v + 1]; ++q) {
          int32_t k = N->vedges[q];
          if (k > j && !seenK[k]) { seenK[k] = 1; candK[nk++] = k; }
        }
    }
    for (int a = 0; a < ni; ++a) {
      int32_t i = candI[a];
      const int32_t* ni_ = O->nodes + O->off[i];
      int si = (int)(O->off[i + 1] - O->off[i]);
      int cij = inter_size(ni_, si, nj, sj);
      if (cij <= 0) continue;
      for (int b = 0; b < nk; ++b) {
        int32_t k = candK[b];
        const int32_t* nk_ = N->nodes + N->off[k];
        int sk = (int)(N->off[k + 1] - N->off[k]);
        int cjk = 0, x = 0, y = 0;
        while (x < sj && y < sk) {  /* j ∩ k, kept for the common part */
          if (nj[x] < nk_[y]) ++x;
          else if (nj[x] > nk_[y]) ++y;
          else { jk[cjk++] = nj[x]; ++x; ++y; }
        }
        if (cjk <= 0) continue;
        int cik = inter_size(ni_, si, nk_, sk);
        if (cik <= 0) continue;
        int gijk = inter_size(ni_, si, jk, cjk);
        out[mochy_bin(si, sj, sk, cij, cjk, cik, gijk)]++;
      }
    }
    for (int a = 0; a < ni; ++a) seenI[candI[a]] = 0;
    for (int b = 0; b < nk; ++b) seenK[candK[b]] = 0;
  }
  free(candI); free(candK); free(seenI); free(seenK); free(jk);
  for (int l = 0; l < 3; ++l) csr_free(&g[l]);
  return 0;
}

int oracle_strict_inc_brute(const int64_t* const off[3], const int32_t* const nodes[3],
                            const int32_t n[3], int64_t out[30]) {
  csr_t g[3];
  for (int i = 0; i < 30; ++i) out[i] = 0;
  if (strict_inc_build(off, nodes, n, g)) return -1;
  int32_t* ij = (int32_t*)malloc(sizeof(int32_t) * 4096);
  for (int32_t i = 0; i < g[0].E; ++i) {
    const int32_t* a = g[0].nodes + g[0].off[i];
    int sa = (int)(g[0].off[i + 1] - g[0].off[i]);
    for (int32_t j = i + 1; j < g[1].E; ++j) {
      const int32_t* b = g[1].nodes + g[1].off[j];
      int sb = (int)(g[1].off[j + 1] - g[1].off[j]);
      int cab = 0, x = 0, y = 0;
      while (x < sa && y < sb) {
        if (a[x] < b[y]) ++x;
        else if (a[x] > b[y]) ++y;
        else { ij[cab++] = a[x]; ++x; ++y; }
      }
      if (!cab) continue;
      for (int32_t k = j + 1; k < g[2].E; ++k) {
        const int32_t* c = g[2].nodes + g[2].off[k];
        int sc = (int)(g[2].off[k + 1] - g[2].off[k]);
        int cbc = inter_size(b, sb, c, sc), cca = inter_size(c, sc, a, sa);
        if (!cbc || !cca) continue;
        out[mochy_bin(sa, sb, sc, cab, cbc, cca, inter_size(ij, cab, c, sc))]++;
      }
    }
  }
  free(ij);
  for (int l = 0; l < 3; ++l) csr_free(&g[l]);
  return 0;
}
