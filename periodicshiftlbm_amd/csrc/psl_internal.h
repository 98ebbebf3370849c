// Library-internal declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/psl.h"
#include "psl_math.h"

namespace psl {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define PSL_HIP_TRY(expr)                                                                       \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) {                                                                     \
      ::psl::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                      \
      return PSL_E_HIP;                                                                         \
    }                                                                                           \
  } while (0)

#define PSL_CHECK_ARG(cond, msg)        \
  do {                                  \
    if (!(cond)) {                      \
      ::psl::set_error(msg);            \
      return PSL_E_ARG;                 \
    }                                   \
  } while (0)

#define PSL_LAUNCH_CHECK()                                                                      \
  do {                                                                                          \
    hipError_t _e = hipGetLastError();                                                          \
    if (_e != hipSuccess) {                                                                     \
      ::psl::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));                 \
      return PSL_E_HIP;                                                                         \
    }                                                                                           \
  } while (0)

// ------------------------------------------------------------- storage types
// FP32 stores f; the 16-bit storages hold d = f - w as raw 16-bit words.
struct StF32 {
  typedef float T;
  static constexpr bool W_OFF = false;
  __device__ __forceinline__ static float load(T x) { return x; }
  __device__ __forceinline__ static T store(float v) { return v; }
};
struct StF16 {
  typedef uint16_t T;
  static constexpr bool W_OFF = true;
  __device__ __forceinline__ static float load(T x) { return __half2float(__ushort_as_half(x)); }
  __device__ __forceinline__ static T store(float v) { return __half_as_ushort(__float2half_rn(v)); }
};
struct StBF16 {
  typedef uint16_t T;
  static constexpr bool W_OFF = true;
  __device__ __forceinline__ static float load(T x) { return __uint_as_float((uint32_t)x << 16); }
  __device__ __forceinline__ static T store(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
};

inline size_t storage_bytes(int storage) { return storage == PSL_F32 ? 4 : 2; }

// 27 shift offsets passed by value: they live in SGPRs inside the kernels.
struct Shift27 {
  int s[27];
};

struct BcTable {
  float v[16];  // 4 slots x (rho, ux, uy, uz)
};

}  // namespace psl

// ------------------------------------------------------------------ handles
namespace psl {
struct MlpNet;
// One workgroup of a fused update launch: a block of 256 bulk cells
// (b == FUSED_BULK, a = block index) or the boundary-list entries [a, b).
struct FusedEntry {
  uint32_t a, b;
};
constexpr uint32_t FUSED_BULK = 0xffffffffu;
// Bulk blocks of at most this many per boundary workgroup in the fused order.
constexpr uint32_t FUSED_SPAN = 32;
// The fused launch order of a grid of n cells with box-face list `list` (cell
// order): each boundary workgroup is placed right before the bulk blocks
// whose cells it covers (at most FUSED_SPAN blocks, at most 256 entries), so
// that with the XCD-contiguous block order both run on one XCD close in time
// and the strided boundary lines are served from that XCD's L2.
std::vector<FusedEntry> fused_order(const std::vector<uint32_t>& list, uint32_t n);
}  // namespace psl

struct psl_grid {
  psl_info info;
  int storage = PSL_F32;
  void* f = nullptr;              // 27 * N storage elements, q-major
  int shift[27] = {0};            // authoritative shifter (host)
  uint8_t* flags = nullptr;       // device, N bytes, or null when all fluid
  std::vector<uint8_t> host_flags;  // user flags before the covered overlay (empty: all fluid)
  uint32_t* bc_list = nullptr;    // device list of box-face boundary cells
  uint32_t n_bc = 0;
  std::vector<psl::FusedEntry> fused_host;  // fused launch order (empty: no list)
  psl::FusedEntry* fused_dev = nullptr;     // device copy
  float* stage_buf = nullptr;       // staging records of the listed cells (28 floats each)
  uint32_t* stage_base = nullptr;   // listed cells below cell 64 w, per wave w
  psl::BcTable bcvals{};
  float4* bc_dev = nullptr;       // device copy of the given-value table (4 x float4)
  float cs = 0.f;
  float* cs_cell = nullptr;       // device, N floats, or null
  int periodic_copies = 0;        // bit 0: x, bit 1: y
  hipStream_t stream = nullptr;
  double* partials = nullptr;     // reduction scratch
  size_t n_partials = 0;
  double* result = nullptr;       // reduction result (device, 4 doubles)
  bool profile = false;           // bracket bulk launches with events
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
  int64_t census[3] = {0, 0, 0};  // bulk, boundary-list, skipped cells
  int tuning = PSL_TUNE_XCD_REMAP;  // bulk-kernel variant (psl_grid_set_tuning / autotune)
  int occupancy = 0;               // bulk-kernel workgroups per CU cap (0: none)
  size_t bulk_lds = 0;             // dynamic LDS a bulk launch reserves for that cap
  const psl_mlp* closure = nullptr;  // learned closure (null: Smagorinsky)
  int closure_every = 1;          // network evaluated every closure_every local updates
  uint64_t closure_count = 0;     // local updates since the closure was set
  uint16_t* closure_y = nullptr;  // held network output per cell (FP16 bits), closure_every > 1
  float* scalar_T = nullptr;      // scalar-transport populations (27 * N FP32, same layout as f)
  float* scalar_given = nullptr;  // prescribed scalar per cell (NaN: free), or null
  float tau_t = 1.f;              // scalar relaxation time
};

// device image of the learned closure of a psl_mlp (psl_closure.hip)
const psl::MlpNet* psl_mlp_device_net(const psl_mlp* m);

struct psl_mask {
  uint8_t* d = nullptr;
  int64_t n = 0;
};

// esotwist (DIAD) grid: fArray(q, cell) at f[q * n + cell], ijk[3][n],
// nbr[7][n] (+i +j +k +ij +ik +jk +ijk), and, on a coarser level of a
// hierarchy, the coarse<->fine interface tables of DIADGridStruct
// (include/types.h:166-171).
struct psl_diad {
  psl_info info;
  int storage = PSL_F32;
  int64_t n = 0;
  void* f = nullptr;
  int* ijk = nullptr;
  int* nbr = nullptr;
  // per cell: bit 0 set when the +i neighbour is the next cell in memory,
  // bits 1, 2, 3 when +ij, +ik, +ijk follow +j, +k, +jk (psl_esotwist.hip)
  uint8_t* adj = nullptr;
  // per coarse-to-fine entry and child: the child's eight slot cells, read
  // by the transfer to the finer level; built for one fine level (key)
  uint32_t* c2f_slots = nullptr;
  uint64_t c2f_slots_key = 0;
  uint8_t* flags = nullptr;
  float4* bc_dev = nullptr;
  float cs = 0.f;
  bool flipper = false;
  hipStream_t stream = nullptr;
  int* c2f_read = nullptr;   // coarse cells
  int* c2f_write = nullptr;  // fine cells
  int64_t n_c2f = 0;
  int* f2c_write = nullptr;  // coarse cells
  int* f2c_read = nullptr;   // fine cells
  int64_t n_f2c = 0;
  int* outlet = nullptr;     // outletCellArray
  int64_t n_outlet = 0;
  int* kmin = nullptr;       // kMinNbrArray (n entries), with the outlet
  int* moving = nullptr;     // moving-bounceback cells
  float* moving_u = nullptr; // their wall velocities (3 per cell)
  int64_t n_moving = 0;
  int* forced = nullptr;     // forced-velocity cells
  float* forced_u = nullptr; // their target velocities (3 per cell)
  int64_t n_forced = 0;
  std::vector<int> moving_host, forced_host;  // host copies of the two lists
  double* partials = nullptr;  // reduction scratch
  size_t n_partials = 0;
  // interface records (psl_esotwist.hip): the local update also writes the
  // post-collision populations of the cells a level transfer reads, one
  // contiguous record per cell, so the transfers read whole lines.  Child
  // records serve the transfer to the coarser level (the children of its
  // fine-to-coarse entries), source records the transfer to the finer level
  // (the sources of this level's coarse-to-fine entries).
  struct Records {
    uint64_t* mask = nullptr;  // per 64 cells: the cells with a record
    uint32_t* base = nullptr;  // per 64 cells: records below them
    void* rec = nullptr;       // ESO_REC_STRIDE storage elements per record
    int64_t n = 0;
    bool fresh = false;        // they hold the last local update's values
  };
  Records child, source;
  int* f2c_child_rec = nullptr;        // (coarser level) 8 per f2c entry: the child's record or -1
  int* c2f_src_rec = nullptr;          // (coarser level) per c2f entry: the source's record or -1
  uint64_t rec_version = 1;            // bumped by every change the records depend on
  uint64_t rec_key = 0;                // versions and neighbours the records were built for
  const psl_diad* rec_fine = nullptr;  // the finer level f2c_child_rec indexes into
};

struct psl_stl {
  std::vector<float> tri;  // 9 floats per triangle (host copy)
  float* d_tri = nullptr;  // device copy
  int count = 0;
  float xmin, ymin, zmin, xmax, ymax, zmax;
  float ox = 0.f, oy = 0.f, oz = 0.f;
};

namespace psl {
// Packed cell coordinates of DIAD cell lists: (k << 42) | (j << 21) | i, so
// that the integer order is the (k, j, i) order of the reference's sortIJK.
constexpr int KEY_BITS = 21;
constexpr uint64_t KEY_MASK = (1ull << KEY_BITS) - 1;
__host__ __device__ inline uint64_t pack_key(int i, int j, int k) {
  return ((uint64_t)k << (2 * KEY_BITS)) | ((uint64_t)j << KEY_BITS) | (uint64_t)i;
}
// implemented in psl_stl.hip: ORs the inside marker of an STL (grid origin go,
// spacing res, box nx x ny x nz) into out[c] for n cells given as packed keys
int stl_mark_keys(const psl_stl* s, float gox, float goy, float goz, float res, int nx, int ny, int nz,
                  const uint64_t* keys, int64_t n, uint8_t* out);
// Restart state (psl_state.hip, psl_esotwist.hip): this header, then the
// storage words of f (27 x n, q-major, as held on the device), then the
// sections the header lists.
struct StateHeader {
  char magic[8];           // "PSLSTATE"
  uint32_t version;        // STATE_VERSION
  uint32_t kind;           // STATE_PS or STATE_DIAD
  int32_t storage;
  int32_t sections;        // STATE_SCALAR | STATE_CLOSURE_Y
  int64_t n;               // cells
  int32_t nx, ny, nz;
  int32_t flipper;         // esotwist parity (DIAD)
  int32_t shift[27];       // shifter (periodic shift)
  int32_t closure_every;   // network evaluated every closure_every updates
  uint64_t closure_count;  // updates since the closure was set
  uint8_t pad[88];
};
static_assert(sizeof(StateHeader) == 256, "state header layout");
constexpr uint32_t STATE_VERSION = 1, STATE_PS = 0, STATE_DIAD = 1;
constexpr int32_t STATE_SCALAR = 1;     // 27 x n FP32 scalar populations follow f
constexpr int32_t STATE_CLOSURE_Y = 2;  // n FP16 held network outputs follow
// a header for kind/storage/n/dims; its validation against a grid on load
void state_header_init(StateHeader& h, uint32_t kind, int storage, int64_t n, const psl_info& info);
int state_header_check(const StateHeader& h, uint32_t kind, int storage, int64_t n, const psl_info& info,
                       int64_t bytes, int64_t expect_bytes);
// implemented in psl_esotwist.hip: the adjacency bits of a DIAD grid from its
// neighbour table (d->nbr set), allocated into d->adj
int diad_build_adjacency(psl_diad* d);
// implemented in psl_esotwist.hip: a version number for a DIAD grid's
// interface records, unique in the process
uint64_t diad_next_version();
// implemented in psl_esotwist.hip: rho, ux, uy, uz per cell (4 x n floats,
// device) on the grid's stream
int diad_macro_device(psl_diad* d, float* out4n);
// implemented in psl_scalar.hip: the scalar value of every cell (n floats,
// device) on the grid's stream
int scalar_field_device(psl_grid* g, float* out);
// implemented in psl_grid.hip
int grid_rebuild_flags(psl_grid* g);
int ensure_partials(psl_grid* g, size_t n);
}  // namespace psl
