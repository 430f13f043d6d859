/*
 * ftr_runtime.hpp -- internals shared by the runtime's translation units
 * (runtime.cpp: problems and plain sampling; learn.cpp: learned proposal).
 */
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <tuple>
#include <string>
#include <vector>

#include "../../../include/feyntrop_amd.h"
#include "../kernels/ftr_abi.hpp"
#include "ftr_host.hpp"

extern const char* ftr_embedded_device_source;
extern const char* ftr_embedded_mlp_source;
extern const char* ftr_embedded_jtable_source;

/* noalloc.hip */
extern "C" hipError_t ftrn_scratch_init(int device, void** ptr);
extern "C" hipError_t ftrn_sum_partials_launch(const double* partials, uint64_t nblocks, int nq, double* tot,
                                               hipStream_t s);

using ftrh::Error;

struct Variant {
  int mode = 0, tmode = 0, block = 256, K = 1;
  int cparams = 0; /* 1: the integration kernel is specialised to the kinematic point */
  int lps = 1;     /* lanes per sample: 2 on the deformed path splits the Jacobian's rows over a lane pair */
};

struct Module {
  hipModule_t mod = nullptr;
  hipFunction_t f_int = nullptr, f_rep = nullptr, f_smp = nullptr;
};

/* concurrent streams of a batch graph's bucket launches */
constexpr int FTR_SIDE_STREAMS = 4;

struct ftr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop;
  std::string err;
  std::map<std::string, Module> modules;
  std::map<std::string, hipModule_t> raw_modules; /* every loaded code object */
  std::map<std::string, Module> spec_modules;     /* the same modules, keyed by the graph spec alone */
  std::map<std::string, std::array<int, 3>> tune_memo; /* tuning file -> (table mode, block, lanes per
                                                        * sample), table mode -1: none */
  /* jump polynomials per (S, 2E-1, log2 block, log2 blocks) and the kernel
   * that chains the block start states */
  std::map<std::tuple<uint64_t, uint64_t, int, int>, std::vector<uint64_t>> jump_cache;
  hipFunction_t f_block_states = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  /* this device's fixed scratch of ftr_integrate_noalloc (noalloc.hip) */
  void* na_scratch = nullptr;
  std::string na_err;
  /* growable device scratch: block states, lane polys, partials */
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  std::vector<double> h_partials;
  /* batched scans */
  void* d_batch = nullptr;
  size_t batch_bytes = 0;
  double* h_pinned = nullptr; /* pinned copy-back buffer of the batch graph */
  size_t pinned_bytes = 0;
  int jtable_mode = 0;        /* 0 auto, 1 host, 2 device (ftr_set_jtable_mode) */
  std::vector<std::string> batch_errors; /* messages of ftr_prepare_batch */
  hipGraphExec_t batch_exec = nullptr;   /* last batch graph, replayed for equal shapes */
  hipStream_t side[FTR_SIDE_STREAMS] = {}; /* the batch graph's concurrent bucket launches */
  hipEvent_t ev_fork = nullptr, ev_join[FTR_SIDE_STREAMS] = {};
  std::string batch_key;
  int batch_replays = 0;
};

struct ftr_problem {
  ftrh::Prep prep;
  ftr_ctx* ctx = nullptr;
  int device = 0;
  Variant var;
  size_t lds = 0;
  Module m;
  ftr::DevParams hprm;
  void* d_slab = nullptr; /* one allocation: parameters and sampler tables */
  void* d_rows_slab = nullptr; /* table mode 3: rows and state records */
  /* problems prepared together share one allocation (d_slab and
   * d_rows_slab then stay null; the last holder frees it) */
  std::shared_ptr<void> d_shared;
  double* d_rows = nullptr;
  ftr::StateRec* d_rec = nullptr;
  int rows_tm = 0; /* placement whose rows d_rows holds (3 or 4; 0: none) */
  ftr::DevParams* d_prm = nullptr;
  double *d_w = nullptr, *d_jr = nullptr, *d_ir = nullptr;
  uint8_t* d_mml = nullptr;
};

/* ------------------------------------------------------------ helpers -- */
#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) throw Error(FTR_ERR_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

/* Geometry + stream positioning for a run of n samples starting at sample
 * `first` of the stream seeded with `seed`.  Writes block states and lane
 * polynomials to the context scratch. */
struct Geometry {
  uint64_t nblocks = 0, S = 0, lanes = 0;
  uint64_t* d_states = nullptr;
  uint64_t* d_polys = nullptr;
  double* d_partials = nullptr;
  double* d_totals = nullptr;   /* nq doubles after the partials */
};

/* kernel source of a problem's variant; `learned` adds the learned-proposal
 * kernels */
std::string full_source(const ftrh::Prep& pr, const Variant& v, bool learned = false);
/* the device parameters of a prepared problem */
ftr::DevParams make_params(const ftrh::Prep& pr);
/* code object for a source (hipRTC, cached on disk) */
std::string code_object(const std::string& src, std::string* path_out);
/* load (once per context) and resolve the kernels named `names` */
hipModule_t load_module_raw(ftr_ctx* c, const std::string& src);
/* Launch geometry and stream positions of n samples from sample `first`:
 * blocks of `block` threads, lps threads per sample slot (each slot draws a
 * contiguous run of S samples); lanes_req sample slots, or 0 to fill the
 * device.  ext (optional): a caller-provided device scratch of ext_bytes
 * instead of the context's growable one; a launch that does not fit is an
 * FTR_ERR_LIMIT error. */
Geometry setup_stream(ftr_problem* p, uint64_t first, uint64_t n, uint64_t seed, uint64_t lanes_req, int block,
                      int nq, hipFunction_t occ_fn, char* ext = nullptr, size_t ext_bytes = 0, int lps = 1);
/* the kernel that writes the block start states (loaded once per context) */
hipFunction_t block_states_fn(ftr_ctx* c);
void ensure_scratch(ftr_ctx* c, size_t bytes);
void prep_from_input(ftrh::Prep& pr, const ftr_input* in);
Variant choose_variant(const ftrh::Prep& pr, size_t lds_limit, size_t mem_limit = (size_t)128 << 30);
size_t variant_lds(const Variant& v, int E);
/* the variant keeps the deformation's second derivatives in LDS: every
 * deformed-path row-table kernel with blocks of at most 256 threads and
 * E <= FTR_DD_LDS_MAX_E (E(E+1)/2 doubles per thread; 110 KiB per block at
 * E = 10), and every two-lanes-per-sample kernel (E(E+1)/2 doubles per lane
 * pair).  FEYNTROP_AMD_DDL=0 keeps them in registers in one-lane kernels. */
int dd_lds_override();
inline bool dd_in_lds(const Variant& v, int E) {
  if (v.mode == 2 && v.lps == 2) return true;
  if (dd_lds_override() == 0) return false;
  return v.mode == 2 && v.tmode >= 3 && v.block <= 256 && E <= FTR_DD_LDS_MAX_E;
}
int fail(ftr_ctx* c, int code, const std::string& msg);

/* the J-table on the device (jtable_gpu.cpp); fills pr.Jr/r/mm/Lt */
void device_jtable(ftr_ctx* c, ftrh::Prep& pr, const ftrh::JtSetup& js, ftrh::JtResult& res);
/* build the row tables of placement tm (table mode 3 or 4) on the device
 * (jtable_gpu.cpp): device_rows allocates (unless the rows of that
 * placement are there) and waits; launch_rows fills rows_region_bytes(tm, E)
 * at d on the context's stream without waiting */
void device_rows(ftr_ctx* c, ftr_problem* p, int tm);
size_t rows_region_bytes(int tm, int E);
void launch_rows(ftr_ctx* c, ftr_problem* p, int tm, char* d);
/* bytes of the table-mode-3 tables of a graph with E edges */
inline size_t rows_bytes(int E) { return ((size_t)1 << E) * ((size_t)8 * E + sizeof(ftr::StateRec)); }
/* table mode 4: folded rows */
inline size_t frows_bytes(int E) { return ((size_t)1 << E) * 8 * (size_t)ftr::frow_stride(E); }
/* device bytes of the row tables of placement tm (3 or 4) */
inline size_t table_rows_bytes(int tm, int E) { return tm == 4 ? frows_bytes(E) : rows_bytes(E); }
/* the table pointers of a problem into a launch */
void fill_tables(ftr::Launch& a, const ftr_problem* p);
/* smallest E for which ftr_prepare builds the J-table on the device in auto mode */
constexpr int FTR_DEVICE_JTABLE_MIN_E = 13;
extern thread_local std::string g_host_err;
