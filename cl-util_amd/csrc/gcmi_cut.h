// Internal host-side state of one grid cut (shared by the plain, batched and
// serving forms).
#pragma once

#include "gcmi_cut_kernels.h"

namespace gcmi {

constexpr int kDefaultSweeps = 4;   // in-tile sweeps per residency
constexpr int kDefaultBfs = 1;      // push rounds per global relabel
constexpr int kRoundBatch = 2;      // BFS / flood rounds launched per host check

enum { CNT_ACT = 0, CNT_BFS = 1, CNT_POS = 2, CNT_FLOOD = 3, CNT_N = 8 };
enum {
  STAT_ROUNDS = 0,
  STAT_RESIDENCIES = 1,
  STAT_GLOBAL_RELABELS = 2,
  STAT_BFS_PASSES = 3,
  STAT_FLOOD_PASSES = 4,
  STAT_SYNCS = 5,
  STAT_ALG_BYTES = 6,
  STAT_TILES = 7,
  STAT_N = 8
};

struct CutState;

// Kernel classes timed by the optional HIP-event profiler.
enum {
  K_PUSH_FIRST = 0,
  K_PUSH = 1,
  K_FOLD = 2,
  K_BFS_INIT = 3,
  K_BFS_ROUND = 4,
  K_FLOOD = 5,
  K_LABELS = 6,
  K_COMPACT = 7,
  K_N = 8
};

// Per-launch HIP events on the cut's stream, summed per kernel class after
// the run's final synchronisation.
struct Prof {
  bool on = false;
  int used = 0;
  int cap = 0;
  hipEvent_t* ev = nullptr;  // 2 per launch
  int* cls = nullptr;
  double ms[K_N];
  int64_t launches[K_N];
  void reset();
  void begin(hipStream_t s, int k);
  void end(hipStream_t s);
  void collect();
  void release();
};

struct CutOps {
  int tile;
  int threads;
  void (*push)(CutState*, int grid, int sweeps, bool first);
  void (*fold)(CutState*, int grid);
  void (*bfs_init)(CutState*, int grid);
  void (*bfs_round)(CutState*, int grid);
  void (*flood)(CutState*, int grid, bool first);
  void (*labels)(CutState*, uint8_t* out);
  void (*untile)(CutState*, const int32_t* src, int32_t* out);
};

struct CutState {
  const CutOps* ops;
  Geo g;
  hipStream_t stream;
  int32_t *e, *h, *cu, *cd, *cl, *cr;  // tile-blocked state
  uint8_t *mask, *reach;               // tile-blocked residual bits / reach
  int32_t* pend;                       // pending edge inflow, 4*T per tile
  int32_t* pflag;                      // tile has pending inflow
  long long* tpos;                     // per tile: sum of max(e, 0)
  int32_t *act, *bflag, *fflag, *haspos;
  int32_t *list, *plist, *alist;       // BFS/flood list, excess tiles, active tiles
  int32_t* counts;                     // device list lengths
  unsigned long long* sums;            // [0] sum min, [1] sum e0+, [2] sum e+
  int32_t* scratch;
  int32_t* in[4];                      // row-major source, sink, right, down
  uint8_t* labels;                     // row-major result
  void* host;                          // pinned readback of the counts
  unsigned long long* host_sums;       // pinned readback of the sums
  size_t device_bytes;
  bool uploaded, solved;
  int64_t flow;
  int64_t stats[STAT_N];
  Prof* prof;
};

const CutOps* cut_ops(int tile);
CutDev cut_dev(CutState* c);
size_t cut_state_bytes(int W, int H, int tile);
int cut_state_init(CutState* c, int W, int H, int tile, hipStream_t stream);
void cut_state_free(CutState* c);
int cut_validate_inputs(const int32_t* cs, const int32_t* ct, const int32_t* right,
                        const int32_t* down, size_t n);
int cut_run(CutState* c, int sweeps, int bfs, int64_t max_rounds, int64_t* flow_out);

}  // namespace gcmi
