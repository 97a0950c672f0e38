/*
 * hmotif.h — C ABI of libhmotif, an MI355X (gfx950) engine for exact hypergraph-motif counting.
 *
 * The ABI takes plain pointers and sizes only. The library owns all device memory; the caller
 * owns host arrays, which are borrowed for the duration of a call. Every entry point returns an
 * hm_status (never exits the process); hm_last_error() gives a message for the calling thread.
 *
 * Hypergraph input format (all entry points): hyperedge e has the nodes
 * nodes[offsets[e] .. offsets[e+1]) — int32 node ids >= 0, offsets int64, offsets[0] == 0.
 * Repeated nodes inside one hyperedge are dropped, as MoCHy's reader does
 * (MoCHy-master/read_data.cpp:34). Hyperedges with identical node sets are distinct hyperedges
 * (MoCHy keeps them) in static counts; in temporal counts they are copies of one static
 * hyperedge (THyMe-main/code/main_thymeP.cpp:164-179).
 *
 * Output numbering:
 *   static  — 30 bins in MoCHy order (MoCHy-master/motif_id.cpp:5-34). MoCHy prints 26 of them
 *             ("motif 1..26" = bins 2,3,5,7,8,...,29; MoCHy-master/main_exact.cpp:133-137);
 *             bins 0,1,4,6 only occur with duplicate hyperedges.
 *   temporal — 96 bins in THyMe order (THyMe-main/code/motif_id.cpp:25-42), bin i = motif i+1.
 */
#ifndef HMOTIF_H
#define HMOTIF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HM_STATIC_BINS 30
#define HM_TEMPORAL_BINS 96
/* Largest hyperedge (in nodes) the counting kernels accept. */
#define HM_MAX_EDGE_SIZE 64

typedef enum hm_status {
  HM_OK = 0,
  HM_ERR_OOM = 1,          /* a device or host allocation failed */
  HM_ERR_OVER_BUDGET = 2,  /* the request does not fit the HBM budget / an index range */
  HM_ERR_BAD_INPUT = 3,    /* malformed arguments */
  HM_ERR_DEVICE = 4,       /* HIP runtime error or no usable gfx950 device */
  HM_ERR_UNSUPPORTED = 5,  /* input outside what the engine handles (e.g. hyperedge > 64 nodes) */
  HM_ERR_BUFFER = 6        /* a caller-provided buffer is too small */
} hm_status;

const char* hm_status_string(hm_status st);
/* Message for the last failing call on this thread ("" when the last call succeeded). */
const char* hm_last_error(void);

/* Venn-id -> bin tables used by the kernels: out[id] = 0-based bin, -1 = not a connected triple.
 * which = 0: static MoCHy bins (MoCHy-master/motif_id.cpp:5-22, minus 1);
 * which = 1: temporal THyMe bins, ids in time order (THyMe-main/code/motif_id.cpp:25-42, minus 1). */
hm_status hm_bin_table(int32_t which, int8_t out[128]);

/* ------------------------------------------------------------------------------------------ */
/* Seeded synthetic generator (replaces dynamicHyperGraph/src/graphGeneration.cpp:10-51,      */
/* which is unseeded). Sizes: power law P(k) ~ k^-alpha on [kmin, kmax]; members: k distinct  */
/* nodes by successive sampling from Zipf(zipf_s) over ranks 1..num_nodes (id = rank - 1).    */
/* ------------------------------------------------------------------------------------------ */
typedef struct hm_gen_params {
  int64_t num_nodes;
  int64_t num_edges;
  double alpha;
  int32_t kmin;
  int32_t kmax;
  double zipf_s;
  uint64_t seed;
} hm_gen_params;

/* Number of incidences (sum of sizes) hm_gen_static will produce for these parameters. */
hm_status hm_gen_static_size(const hm_gen_params* p, int64_t* num_incidences);
/* offsets: num_edges + 1 entries; nodes: nodes_capacity >= hm_gen_static_size(). */
hm_status hm_gen_static(const hm_gen_params* p, int64_t* offsets, int32_t* nodes,
                        int64_t nodes_capacity);
/* n strictly increasing int64 timestamps (microseconds): singleton gaps Exp(mean_gap_us);
 * round(n*burst_fraction/burst_len) bursts of burst_len edges with Exp(burst_gap_us) gaps;
 * every gap is rounded and at least 1 us. */
hm_status hm_gen_timestamps(int64_t n, double mean_gap_us, double burst_fraction,
                            int32_t burst_len, double burst_gap_us, uint64_t seed, int64_t* times);
/* k distinct indices drawn uniformly from [0, n) (seeded partial Fisher-Yates). */
hm_status hm_gen_pick(int64_t n, int64_t k, uint64_t seed, int64_t* idx);

/* ------------------------------------------------------------------------------------------ */
/* Device runtime                                                                             */
/* ------------------------------------------------------------------------------------------ */
/* Selects the HIP device used by handles created afterwards on this thread (default 0) and
 * checks that it is a gfx950 part. */
hm_status hm_init(int32_t device);

/* ------------------------------------------------------------------------------------------ */
/* Static hypergraph handle: flat sorted H2V / V2H arrays resident in HBM.                    */
/* Replaces CBSTOperations::construct (dynamicHyperGraph/include/structure.hpp:45-64) and     */
/* the host projection hyperedgeAdjacency (dynamicHyperGraph/utils/utils.cpp:399-427).        */
/* ------------------------------------------------------------------------------------------ */
typedef struct hm_graph hm_graph;

/* num_nodes <= 0 means max node id + 1 (MoCHy's V). Copies the CSR to the device. */
hm_status hm_graph_create(const int64_t* offsets, const int32_t* nodes, int32_t num_edges,
                          int32_t num_nodes, hm_graph** out);
hm_status hm_graph_destroy(hm_graph* g);

/* Exact static h-motif counts, 30 MoCHy bins.
 * Replaces computeMotifCounts (dynamicHyperGraph/include/motif.hpp:10-13,
 * dynamicHyperGraph/src/HMotifCount.cu:97-125), which counted closed triples only into int32
 * and printed them; the count here follows MoCHy-E (MoCHy-master/main_exact_par.cpp:122-165). */
hm_status hm_count_static(hm_graph* g, uint64_t out[HM_STATIC_BINS]);

/* Static count that allocates nothing. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same handle, output and counts as hm_count_static; its scratch is a static device
 * array of HM_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on that GPU until the
 * process ends, instead of the handle's own buffers. Limits: the hyperedge ids must fit one LDS
 * table (hm_count_static's automatic one-table path, no forced tile), and ~65 B per hyperedge plus
 * 4 B per directed projection pair must fit the array (the 5,000-hyperedge static workload needs
 * ~43 MB); otherwise HM_ERR_OVER_BUDGET (or HM_ERR_UNSUPPORTED for a forced tile) and no count.
 * The call never calls hipMalloc, hipHostMalloc or hipFree and never autotunes (it uses the
 * handle's triangle form, else the one tuned for its shape, else 8 waves x 16 copies). Calls on one
 * GPU share its array: the library runs them one at a time, and each drains its stream before it
 * returns. */
#define HM_NOALLOC_SCRATCH_BYTES (64ll << 20)
hm_status hm_count_static_noalloc(hm_graph* g, uint64_t out[HM_STATIC_BINS]);

/* Key tile of hm_count_static and of the delta of hm_graph_update / hm_count_delta on g.
 * 0 (the default): one LDS mask table over all hyperedge ids when they fit (about 28k ids for
 * the static count and the delta with 32-bit masks), tiles as large as LDS allows otherwise; > 0: tiles of
 * exactly tile_ids ids forced at any size (tests and tuning). Results do not depend on the tile.
 * No counterpart in the reference, whose kernel keeps no per-hyperedge table
 * (dynamicHyperGraph/src/HMotifCount.cu:16-95). */
hm_status hm_graph_set_tile(hm_graph* g, int32_t tile_ids);

/* Form of the one-table triangle kernel of hm_count_static on g (waves per centre workgroup and
 * int32 LDS histogram copies per wave). waves = 0 (default): the form autotuned per shape (mask
 * width, log2 of the hyperedge count) on the first count of that shape in the process — each of
 * the 4 forms (4 or 8 waves x 8 or 16 copies) run once, their bins checked equal, the fastest
 * kept; waves = -1: as 0, but tune the shape again on the next count; otherwise waves 4 or 8 and
 * copies 8 or 16. Counts are identical in every form. No counterpart in the reference, whose
 * kernel runs one thread per hyperedge (dynamicHyperGraph/src/HMotifCount.cu:16-95). */
hm_status hm_graph_set_tri_form(hm_graph* g, int32_t waves, int32_t copies);
/* Form used by the last one-table count of g (0, 0 after a tiled count) and whether that count
 * ran the autotuner. */
hm_status hm_graph_get_tri_form(hm_graph* g, int32_t* waves, int32_t* copies, int32_t* tuned);

/* on != 0: after an eager hm_count_static, the whole count of g (every launch, no host wait) is
 * captured as a hipGraph, and later counts of the unchanged graph replay it — one graph launch
 * and one wait instead of ~15 launches and two waits. Any update or tile change drops the capture
 * (the next count runs eagerly and captures again). A capture that contains a memset node is
 * never replayed. Counts are identical either way. No counterpart in the reference, which
 * launches and synchronizes per kernel (dynamicHyperGraph/src/HMotifCount.cu:97-125). */
hm_status hm_graph_set_replay(hm_graph* g, int32_t on);
/* captured: a capture of the current graph is held; last_replayed: the last count replayed it. */
hm_status hm_graph_replay_state(hm_graph* g, int32_t* captured, int32_t* last_replayed);

/* Graph statistics computed on the device by the last hm_count_static call on g. */
typedef struct hm_graph_stats {
  int64_t num_edges;          /* E */
  int64_t num_nodes;          /* V */
  int64_t num_incidences;     /* I */
  int64_t proj_pairs;         /* P = sum_a d_a (directed projected pairs) */
  int64_t max_proj_degree;    /* max_a d_a */
  int64_t oriented_row_sum;   /* sum_a sum_{b in N(a), b > a} d_b */
  int64_t wedges;             /* sum_a C(d_a, 2) */
  int64_t triangles;          /* closed triples (including the 4 hidden bins) */
  int64_t max_node_degree;    /* max_v |V2H(v)| */
} hm_graph_stats;
hm_status hm_graph_get_stats(hm_graph* g, hm_graph_stats* out);

/* Hub-aware static count (no counterpart in the reference, whose counters enumerate every triple:
 * MoCHy-master/main_exact_par.cpp:122-165, dynamicHyperGraph/src/HMotifCount.cu:16-95).
 * hubs = 0 (default): hm_count_static enumerates the triangles of the full projection; hubs in
 * 1..23: the `hubs` most frequent nodes are hubs, and the count combines hyperedge classes (hub
 * pattern, light part non-empty), the types of the hyperedge pairs that share a light node, the
 * pairs of light neighbours of every hyperedge grouped by key, and the light triangles, without
 * enumerating the cliques that hubs put into the projection; hubs = -1: the hub count is chosen by
 * a cost model (and lowered by a quarter while some hyperedge has more distinct neighbour keys
 * than the largest histogram holds — 7,168 in LDS, then up to 917,504 in global memory; if even
 * one hub overflows, the triangle path counts). The 30 bins are identical
 * either way. A forced hub count whose histograms overflow fails with HM_ERR_UNSUPPORTED. After a
 * hub-aware count hm_graph_stats holds the projection statistics computed from the projected
 * degrees alone (no rows: upper-row totals are not kept), by the first hm_graph_get_stats after
 * the count (they are not part of the bins, so the count itself does not compute them).
 * With hubs != 0, hm_graph_update's delta is hub-aware too: the class, pair-type and light-wedge
 * parts of the graph before and after the batch, and the light triangles that hold a changed
 * hyperedge (deleted ones before the batch, inserted ones after it), with one hub set for both
 * graphs (chosen on the graph before; the automatic hub count, at most 12 for a delta, is kept
 * across updates). If the
 * graph after the batch overflows a neighbour-key histogram with that hub set, the batch stays
 * applied and the update fails (recount the graph). */
hm_status hm_graph_set_hubs(hm_graph* g, int32_t hubs);
typedef struct hm_hub_stats {
  int32_t hubs;           /* hubs of the last hub-aware count on g (0: it counted by triangles) */
  int64_t classes;        /* distinct (hub pattern, [light part non-empty]) classes */
  int64_t light_pairs;    /* hyperedge pairs sharing a light node (pairs of the light projection) */
  int64_t pair_types;     /* distinct pair types among the light pairs */
  int64_t light_triangles; /* hyperedge triples whose three pairs share light nodes */
  double ms_classes;      /* class keys and their sort */
  double ms_wedges;       /* light neighbours, pair-type keys and light wedges */
  double ms_types;        /* class triples, pair-type sort and pair types x classes */
  double ms_triangles;    /* light projection rows and light triangles */
  double ms_tri_kernel;   /* the light-triangle kernel alone (k_hub_tri or k_hub_tri_tiled) */
  int64_t big_centres;    /* hyperedges whose neighbour keys outgrew the LDS histogram (7,168 keys),
                             counted with global-memory ones (57,344, then 229,376, then 917,504 keys) */
} hm_hub_stats;
hm_status hm_graph_get_hub_stats(hm_graph* g, hm_hub_stats* out);

/* One part of a hub-aware count with `hubs` hubs (1..23), so that a long count can be split over
 * calls, processes or devices (no counterpart in the reference, whose counters run one process per
 * graph: MoCHy-master/main_exact_par.cpp:122-165). The light triangles are enumerated once each, at
 * their centre of lowest light rank (ascending light degree, ties by id: the same order in every
 * process holding the same graph); this call enumerates those of the centres with ranks in
 * [centre_begin, centre_end) (centre_end = -1: up to the number of hyperedges) and, with
 * other_parts = 1, adds the class, pair-type and light-wedge parts, which involve every hyperedge.
 * The bins of calls whose ranges partition [0, E), other_parts set in exactly one, sum (mod 2^64)
 * to hm_count_static's with hm_graph_set_hubs(g, hubs). *light_triangles (may be NULL) receives the
 * light triangles of the range. With other_parts = 1, hm_graph_get_stats then holds the
 * projection statistics (wedges = sum_a C(d_a, 2)) with triangles = -1. A neighbour-key histogram
 * overflow (other_parts = 1) returns HM_ERR_UNSUPPORTED. seconds > 0 bounds the call: the range's
 * light triangles are enumerated in launches of at most 1/256 of the centres, and none starts once
 * `seconds` have passed since the call began; *centre_done (may be NULL) receives the end of the
 * ranks enumerated (the range's end unless the time ran out), where the next part begins. */
hm_status hm_graph_count_hub_part(hm_graph* g, int32_t hubs, int64_t centre_begin, int64_t centre_end,
                                  int32_t other_parts, double seconds, uint64_t out[HM_STATIC_BINS],
                                  int64_t* light_triangles, int64_t* centre_done);

/* Kernel timing of the last hm_count_static on g, measured with HIP events on the handle's
 * stream: names are "project" and "triangles"; "store" is the store-update part of the last eager
 * hm_graph_update (0 after a captured one); ms receives the duration of that phase. */
hm_status hm_graph_phase_ms(hm_graph* g, const char* phase, double* ms);

/* Current sizes of g's store: hyperedge ids (deleted ones included, as empty rows), nodes, and
 * incidences. No counterpart in the reference, whose store prints its flat arrays instead
 * (dynamicHyperGraph/utils/utils.cpp:434-435). */
hm_status hm_graph_get_sizes(hm_graph* g, int32_t* num_edges, int32_t* num_nodes, int64_t* num_incidences);
/* Copies g's store to host arrays: H2V offsets[num_edges + 1] and nodes[num_incidences] (rows in
 * hyperedge id order, nodes ascending, a deleted hyperedge an empty row) and V2H
 * voffsets[num_nodes + 1] and vedges[num_incidences] (hyperedge ids ascending per node). Any
 * pointer may be null to skip that array; sizes from hm_graph_get_sizes. Waits for g's stream. */
hm_status hm_graph_read_csr(hm_graph* g, int64_t* offsets, int32_t* nodes, int64_t* voffsets,
                            int32_t* vedges);

/* ------------------------------------------------------------------------------------------ */
/* Dynamic hypergraphs                                                                        */
/* ------------------------------------------------------------------------------------------ */
/* Applies one batch in place on the device: deletes the live hyperedges del_ids (their rows
 * become empty) and inserts n_ins hyperedges given as a CSR (ins_offsets[0] == 0), which receive
 * the fresh ids num_edges, num_edges + 1, ... written to assigned_ids (ids are never reused).
 * If delta_out is not NULL it receives the exact change of the 30 static bins,
 * counts(after) - counts(before), computed incrementally from the changed hyperedges only.
 * Replaces the reference's CBSTOperations::insert / erase and fillCBST / unfillCBST
 * (dynamicHyperGraph/include/structure.hpp:39-55, structure/operations.cu:82-453) as driven by
 * applyIncrementalUpdates (dynamicHyperGraph/src/main.cu:238-250), plus its delta count. */
hm_status hm_graph_update(hm_graph* g, const int32_t* del_ids, int32_t n_del,
                          const int64_t* ins_offsets, const int32_t* ins_nodes, int32_t n_ins,
                          int32_t* assigned_ids, int64_t delta_out[HM_STATIC_BINS]);

/* What the last hm_graph_update on g that computed a delta touched, and where its time went.
 * HM_ERR_BAD_INPUT if the last update had no delta. No counterpart in the reference, whose anchor
 * rule does not separate removed and added triples (dynamicHyperGraph/src/HMotifCountUpdate.cu:151-226). */
typedef struct hm_update_stats {
  int64_t removed;      /* triples that contained a deleted hyperedge before the batch (hub-aware
                           delta: the light triangles among them) */
  int64_t added;        /* triples that contain an inserted hyperedge after it (delta sums to added - removed;
                           hub-aware delta: the light triangles among them) */
  int64_t row_entries;  /* projection-row entries scanned by the delta kernel, both phases (-1: hub-aware) */
  double ms_rows;       /* projection rows of the graph after the batch, and before it unless the
                           rows built by the previous update are still held (HIP events) */
  double ms_delta;      /* delta kernel, both phases (HIP events); hub-aware delta: its parts on
                           both graphs (host time, ms_rows 0) */
  double ms_store;      /* store update: H2V / V2H rebuild, stream time including its host wait */
  int32_t captured;     /* 1: the update ran as one captured graph (hm_graph_set_update_capture);
                           ms_delta then holds the whole launch, ms_rows and ms_store 0 */
  int32_t rows_overflow; /* 1: its rows after the batch outgrew the plan and the addition phase
                           ran eagerly (its row build and delta added to ms_rows / ms_delta) */
} hm_update_stats;
hm_status hm_graph_update_stats(hm_graph* g, hm_update_stats* out);

/* on != 0: every later hm_graph_update of g that asks for a delta (one id table: no forced tile,
 * and the ids after the batch fit one delta table; hyperedges of at most 64 nodes) runs as ONE
 * captured hipGraph — removal delta over the rows held from the previous update, store update
 * (H2V / V2H rebuild), projection rows of the graph after the batch, addition delta — with the
 * batch staged in pinned memory: one copy in, one graph launch, one copy out, one wait, instead
 * of ~30 launches and three host waits. Every buffer is sized to a plan with headroom (ids,
 * nodes, incidences, batch sizes, row entries); a batch that does not fit grows the plan and
 * captures again (the store and the held rows are kept). Rows after the batch that outgrow the
 * plan are caught on the device: the addition phase is skipped and rerun eagerly with larger
 * buffers. A plan that cannot be made (or a capture holding a memset node) leaves g's updates
 * eager. Deltas, stores and ids are identical either way. on = 0 drops the captured graph. No
 * counterpart in the reference, whose update rebuilds both snapshots on the host and launches
 * per step (dynamicHyperGraph/src/main.cu:332-353). */
hm_status hm_graph_set_update_capture(hm_graph* g, int32_t on);
typedef struct hm_update_capture_state {
  int32_t on;               /* capture requested on g */
  int32_t disabled;         /* a plan could not be made or replayed: g's updates run eagerly */
  int32_t captured;         /* an instantiated update graph is held */
  int32_t captures;         /* update graphs captured on g so far (first plan and each regrowth) */
  int32_t last_captured;    /* the last update with a delta ran as the captured graph */
  int32_t last_overflow;    /* ... and its rows outgrew the plan (addition phase rerun eagerly) */
  int64_t id_capacity;      /* hyperedge ids the plan holds (deleted ids keep theirs) */
  int64_t node_capacity;    /* node ids */
  int64_t incidence_capacity;
  int64_t row_capacity;     /* full projection-row entries */
  int32_t batch_deletes;    /* deleted ids per batch */
  int32_t batch_inserts;    /* inserted hyperedges per batch */
  int64_t batch_incidences; /* inserted incidences per batch */
} hm_update_capture_state;
hm_status hm_graph_update_capture_state(hm_graph* g, hm_update_capture_state* out);

/* Exact delta between two snapshots of one id space, counts(new_g) - counts(old_g), given the
 * ids deleted from old_g and the ids inserted into new_g (an id may be deleted and re-inserted;
 * all other hyperedges must be identical in both). Same arguments as computeMotifCountsDelta
 * (dynamicHyperGraph/include/motif_update.hpp:13-19), with 0-based ids and int64 output; the
 * reference's anchor rule (dynamicHyperGraph/src/HMotifCountUpdate.cu:99-226) is replaced by
 * charging each changed triple to its smallest changed id. */
hm_status hm_count_delta(hm_graph* old_g, hm_graph* new_g, const int32_t* del_ids, int32_t n_del,
                         const int32_t* ins_ids, int32_t n_ins, int64_t delta_out[HM_STATIC_BINS]);

/* ------------------------------------------------------------------------------------------ */
/* Three-snapshot strict-increasing count (the reference's temporal-window interface)         */
/* ------------------------------------------------------------------------------------------ */
/* Replaces computeTemporalMotifCountsStrictInc (temporalDynamicHMotif/include/temporal_count.hpp:8-9,
 * temporalDynamicHMotif/src/temporal_count.cu:91-217). Layer 0 = older, 1 = middle, 2 = newest
 * snapshot, each a CSR over its row ids (row r is the reference's hyperedge id r + 1; an empty row
 * is an absent id). Counts the triples (i, j, k), i < j < k, of older row i, middle row j and newest
 * row k that pairwise overlap, into the 30 static bins computed from the three rows. Unlike the
 * reference kernel, which keeps at most 1,024 candidates on each side of j, no candidate is dropped.
 * This is not THyMe's temporal h-motif count (hm_count_temporal). */
hm_status hm_count_strict_inc(const int64_t* const offsets[3], const int32_t* const nodes[3],
                              const int32_t num_rows[3], uint64_t out[HM_STATIC_BINS]);

/* ------------------------------------------------------------------------------------------ */
/* Temporal hypergraph handle: temporal hyperedges (node set, int64 timestamp) in time order  */
/* (ties by input order), resident in HBM.                                                    */
/* Replaces the 3-layer ring buffer TemporalHypergraphIndex /                                 */
/* TemporalCBSTOperations (temporalDynamicHMotif/include/temporal_adjacency.hpp:9-18,         */
/* temporal_structure.hpp:24-78) and the incremental projection of THyMe+                     */
/* (THyMe-main/code/main_thymeP.cpp:54-115).                                                  */
/* ------------------------------------------------------------------------------------------ */
typedef struct hm_tgraph hm_tgraph;

hm_status hm_tgraph_create(const int64_t* offsets, const int32_t* nodes, const int64_t* times,
                           int32_t num_edges, hm_tgraph** out);
hm_status hm_tgraph_destroy(hm_tgraph* g);

/* Exact temporal h-motif counts within window delta (same unit as the timestamps), 96 THyMe
 * bins: every triple of temporal hyperedges with t_max - t_min <= delta and at least two
 * overlapping pairs, classified by its Venn pattern in time order.
 * Replaces computeTemporalMotifCountsStrictInc (temporalDynamicHMotif/include/
 * temporal_count.hpp:10-11), which counted a different, time-free 30-bin pattern; the count here
 * is THyMe+'s (THyMe-main/code/main_thymeP.cpp:219-367). Equal to THyMe+ whenever timestamps are
 * distinct (with ties THyMe's result depends on random tie-breaking; here ties follow input order). */
hm_status hm_count_temporal(hm_tgraph* g, int64_t delta, uint64_t out[HM_TEMPORAL_BINS]);

/* Position tile of hm_count_temporal on g. 0 (the default): one LDS table per δ-window when it
 * fits (windows up to ~19k temporal hyperedges), position tiles as large as LDS allows otherwise;
 * > 0: tiles of exactly tile_positions positions forced at any window size (tests and tuning).
 * The counts do not depend on the tile. No counterpart in the reference (THyMe+ keeps the window
 * in host hash maps, THyMe-main/code/main_thymeP.cpp:54-115). */
hm_status hm_tgraph_set_tile(hm_tgraph* g, int32_t tile_positions);

typedef struct hm_temporal_stats {
  int64_t num_edges;        /* temporal hyperedges */
  int64_t num_incidences;
  int64_t proj_pairs;       /* overlapping pairs of temporal hyperedges within delta */
  int64_t max_window;       /* max hyperedges strictly older than an anchor inside its window */
  int64_t neighbour_pairs;  /* sum over anchors of C(#window neighbours, 2) */
  int64_t instances;        /* sum of the 96 bins */
  int64_t oriented_row_sum; /* sum over later-overlapping pairs (x, y) of y's window degree */
  int64_t num_nodes;        /* max node id + 1 */
} hm_temporal_stats;
hm_status hm_tgraph_get_stats(hm_tgraph* g, hm_temporal_stats* out);
/* Phases "project" (window bounds + temporal projection rows), "count" (k_tcount or
 * k_tcount_tiled) and "wedge" (k_twedge, one-table counts only: the open triples whose earliest
 * edge precedes their middle edge, counted per middle edge). */
hm_status hm_tgraph_phase_ms(hm_tgraph* g, const char* phase, double* ms);

/* ------------------------------------------------------------------------------------------ */
/* Batched static counts: many graphs in one launch sequence.                                 */
/* No counterpart in the reference (it counts one graph per call, HMotifCount.cu:97-125).     */
/* ------------------------------------------------------------------------------------------ */
typedef struct hm_batch hm_batch; /* scratch + stream + tuned forms; one caller thread at a time */

typedef struct hm_batch_stats {
  int32_t graphs_batched; /* graphs counted by the batched kernels */
  int32_t graphs_alone;   /* graphs larger than one LDS table, counted one by one (tiled) */
  int32_t waves;          /* triangle form used: waves per centre (4 or 8), */
  int32_t copies;         /* LDS histogram copies per wave (8 or 16), */
  int32_t persistent;     /* 1 = persistent grid fed by a device work queue, 0 = one group per centre */
  int32_t order;          /* centre order: 0 graph by graph, 1 chunks of graphs with up to 64 MB of
                             projection rows, 2 all centres of the batch by work */
  int32_t tuned;          /* 1 if this call ran the autotuner (each run checked equal) */
  int64_t hyperedges;     /* hyperedges counted by the batched kernels */
  double ms_prepare;      /* device time before the triangle kernel */
  double ms_triangles;    /* triangle kernel (the fastest form when this call tuned) */
} hm_batch_stats;

hm_status hm_batch_create(hm_batch** out);
hm_status hm_batch_destroy(hm_batch* b);
/* Exact static counts of n graphs, out[30 * k + bin] for graphs[k], identical to
 * hm_count_static on each. The graphs must not be used by another call meanwhile; their
 * statistics (hm_graph_get_stats) are not refreshed. The triangle kernel's form and centre order
 * are autotuned per batch shape (mask width, log2 of the total and of the largest hyperedge
 * count) on the first batch of that shape — the three orders, then the eight forms, every run
 * checked equal — unless hm_batch_set_form fixed them. */
hm_status hm_batch_count_static(hm_batch* b, hm_graph* const* graphs, int32_t n, uint64_t* out);
/* waves = 0: autotuned (default); otherwise waves 4 or 8, copies 8 or 16, persistent 0 or 1,
 * order 0, 1 or 2 (see hm_batch_stats). */
hm_status hm_batch_set_form(hm_batch* b, int32_t waves, int32_t copies, int32_t persistent,
                            int32_t order);
hm_status hm_batch_get_stats(hm_batch* b, hm_batch_stats* out);

/* ------------------------------------------------------------------------------------------ */
/* Serving: many device-resident hypergraphs, requests admitted and run as slots free up.     */
/* No counterpart in the reference (its drivers run one graph, one request, one stream).      */
/* ------------------------------------------------------------------------------------------ */
typedef struct hm_server hm_server;

#define HM_REQ_COUNT 0  /* exact static count of the graph as of this request */
#define HM_REQ_UPDATE 1 /* one batch (deletes + inserts); result = delta of the 30 bins */

typedef struct hm_server_config {
  int64_t hbm_budget_bytes; /* graphs are admitted while their estimated footprint fits */
  int32_t slots;            /* requests running at once (worker threads, one stream per graph) */
  /* 0: every count request counts the graph on the device.
   * 1: after a graph's first count, its 30 counts are kept current by adding the exact delta of
   *    each later update (the delta every update computes anyway), and a count request on a
   *    graph with no request ahead of it is answered from them without a launch. The counts are
   *    identical either way; an update that fails drops the kept counts until the next count. */
  int32_t maintain_counts;
  /* 0 or 1: each count runs on its own. n > 1: a worker whose next request is a count of a small
   * graph (its last count took under 1 ms) also takes other ready small graphs whose next request
   * is a count, up to n graphs and about 4 ms of expected work, and counts them with one
   * hm_batch_count_static call (counts identical to the one-graph count). */
  int32_t batch_graphs;
  /* 1: every graph's updates run as one captured hipGraph each (hm_graph_set_update_capture):
   * one launch and one wait per update. Deltas identical either way. */
  int32_t capture_updates;
  /* n > 0: a graph added with at least n hyperedges is counted and updated hub-aware
   * (hm_graph_set_hubs(g, -1): its counts and deltas by hyperedge classes, pair types, light
   * wedges and light triangles; never in a batched count; the captured form is not used for its
   * updates); 0: every graph by triangles. Counts and deltas identical either way. */
  int32_t hub_min_edges;
} hm_server_config;

typedef struct hm_result {
  int32_t status;                /* hm_status of the request */
  int32_t kind;                  /* HM_REQ_COUNT or HM_REQ_UPDATE */
  int64_t counts[HM_STATIC_BINS];/* counts (COUNT) or counts(after) - counts(before) (UPDATE) */
  double queued_ms;              /* submit -> start */
  double run_ms;                 /* start -> finish */
  double latency_ms;             /* submit -> finish */
  int32_t first_new_id;          /* UPDATE: id of the first inserted hyperedge (-1 if none) */
  int32_t n_new;                 /* UPDATE: number of inserted hyperedges (ids are consecutive) */
} hm_result;

typedef struct hm_server_stats {
  int64_t graphs;
  int64_t reserved_bytes;
  int64_t budget_bytes;
  int64_t completed;
  int64_t coalesced;  /* count requests answered by another request's count */
  int64_t maintained; /* count requests answered from kept counts (maintain_counts = 1) */
  int64_t batched_counts; /* graph counts run inside batched launches (batch_graphs > 1) */
  int64_t batch_launches; /* batched count calls with more than one graph */
} hm_server_stats;

hm_status hm_server_create(const hm_server_config* cfg, hm_server** out);
/* Waits for the workers to finish the queued requests and frees every graph. */
hm_status hm_server_destroy(hm_server* s);
/* HM_ERR_OVER_BUDGET if the graph's estimated device footprint does not fit the budget. After every
 * request the server charges what each graph's buffers hold (they grow with updates), so
 * reserved_bytes follows the device memory in use; while it exceeds the budget, new graphs are
 * refused and update requests fail with HM_ERR_OVER_BUDGET (the graph unchanged; counts still
 * run) until a graph is removed. */
hm_status hm_server_add_graph(hm_server* s, const int64_t* offsets, const int32_t* nodes,
                              int32_t num_edges, int32_t* graph_id);
hm_status hm_server_remove_graph(hm_server* s, int32_t graph_id);
/* Non-blocking; requests on one graph run in submission order. */
hm_status hm_server_submit_count(hm_server* s, int32_t graph_id, uint64_t* ticket);
hm_status hm_server_submit_update(hm_server* s, int32_t graph_id, const int32_t* del_ids,
                                  int32_t n_del, const int64_t* ins_offsets,
                                  const int32_t* ins_nodes, int32_t n_ins, uint64_t* ticket);
/* Blocks until the request is done; each ticket can be waited for (or polled to completion) once. */
hm_status hm_server_wait(hm_server* s, uint64_t ticket, hm_result* out);
/* Non-blocking: *done = 0 while the request runs (out untouched, HM_OK); once it is done,
 * *done = 1 and out receives the result exactly as hm_server_wait would return it (the ticket
 * is then consumed). The reference has no request interface; this is the hm_poll of the
 * serving contract next to hm_server_submit_* (submit) and hm_server_wait. */
hm_status hm_server_poll(hm_server* s, uint64_t ticket, int32_t* done, hm_result* out);
hm_status hm_server_get_stats(hm_server* s, hm_server_stats* out);

/* ------------------------------------------------------------------------------------------ */
/* Calibration. No counterpart in the reference.                                               */
/* ------------------------------------------------------------------------------------------ */
/* Runs an integer-VALU-only kernel (8 workgroups of 4 waves per CU, `iters` loop trips of 32
 * independent xor/shift/add steps in the first `lanes` (1..64) lanes of every wave) on the current
 * device and returns the duration of its second launch in *ms. Under rocprofv3 counters its
 * SQ_INSTS_VALU per cycle is the VALU issue peak that the counting kernels' own issue rate is
 * compared with (bench.py --valu-probe); lanes < 64 measures issue with part of EXEC cleared. */
hm_status hm_probe_valu(int32_t iters, int32_t lanes, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* HMOTIF_H */
