/* C ABI of the MI355X distributed-FFT library (libdfft.so).
 *
 * One process drives P "virtual ranks" that share one GPU's HBM: every
 * entry point that takes per-rank data takes a table of P device pointers,
 * one local array per rank, laid out exactly as the reference's per-rank
 * Kokkos::View (LayoutRight unless stated).  Only plain C types cross this
 * boundary; `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Every function returns a status (DFFT_SUCCESS = 0).  On failure
 * dfft_last_error() returns the message of the calling thread's last error;
 * the codes mirror the reference's std::runtime_error throw sites (cited
 * per function).
 */
#ifndef DFFT_H
#define DFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define DFFT_SUCCESS 0
#define DFFT_ERROR_INVALID_TOPOLOGY 1 /* Plan.hpp:62-68, Topologies.hpp throws */
#define DFFT_ERROR_INVALID_EXTENTS 2  /* MPI_Extents.hpp:322-417 */
#define DFFT_ERROR_MISMATCH 3         /* InternalPlan.hpp:109-142, Plan.hpp:212-238 */
#define DFFT_ERROR_ALIAS 4            /* TransBlock.hpp:120-186 */
#define DFFT_ERROR_UNSUPPORTED 5      /* Plan.hpp:85-112 (Brick/Empty), sizes beyond kernels */
#define DFFT_ERROR_HIP 6              /* HIP runtime failure */
#define DFFT_ERROR_INVALID_ARGUMENT 7 /* null pointers, bad enums, rank out of range */
#define DFFT_ERROR_OUT_OF_MEMORY 8

/* ---- enums ------------------------------------------------------------- */
#define DFFT_C2C 0 /* complex -> complex */
#define DFFT_R2C 1 /* real -> complex forward; its backward is C2R */

#define DFFT_FP32 0
#define DFFT_FP64 1

#define DFFT_LAYOUT_RIGHT 0
#define DFFT_LAYOUT_LEFT 1

/* KokkosFFT::Normalization; default of the reference Plan is backward (Plan.hpp:159) */
#define DFFT_NORM_BACKWARD 0
#define DFFT_NORM_FORWARD 1
#define DFFT_NORM_ORTHO 2
#define DFFT_NORM_NONE 3

/* KokkosFFT::Direction */
#define DFFT_FORWARD 0
#define DFFT_BACKWARD 1

/* execution strategies of a plan */
#define DFFT_EXEC_FUSED 0     /* exchange fused into the FFT passes (default) */
#define DFFT_EXEC_BLOCKWALK 1 /* the reference's block sequence: FFT blocks + exchanges */
#define DFFT_EXEC_VENDOR 2    /* rocFFT on the gathered global array (the TplPlan role, TplPlan.hpp:25-181);
                                 a comparator: no adjoint, plan name "rocFFTPlan" */

#define DFFT_MAX_RANK 4
#define DFFT_MAX_VRANKS 64

typedef struct dfft_plan dfft_plan;

/* Plan description: what the reference's Plan constructor receives
 * (Plan.hpp:154-181) with the per-rank views replaced by the global extents
 * the reference reconstructs with MPI_Allgather (MPI_Extents.hpp:114-129). */
typedef struct dfft_desc {
  int rank;                               /* view rank, 2..4 */
  int fft_rank;                           /* number of transformed axes, 1..3 */
  int axes[3];                            /* transform axes (negative allowed) */
  int transform;                          /* DFFT_C2C or DFFT_R2C */
  int precision;                          /* DFFT_FP32 or DFFT_FP64 */
  int layout;                             /* view layout, DFFT_LAYOUT_RIGHT */
  size_t in_extents[DFFT_MAX_RANK];       /* global input extents (real for R2C) */
  size_t out_extents[DFFT_MAX_RANK];      /* global output extents */
  size_t in_topology[DFFT_MAX_RANK];
  size_t out_topology[DFFT_MAX_RANK];
  int in_topology_layout;                 /* Topology<..., LayoutRight|LayoutLeft> */
  int out_topology_layout;
  int norm;                               /* DFFT_NORM_* */
  int nranks;                             /* P: must equal the topology size */
  int exec_mode;                          /* DFFT_EXEC_* */
} dfft_desc;

/* BlockInfo<DIM> (Types.hpp:61-154) as a POD; arrays hold `rank` entries
 * (`buffer_extents` rank+1). */
typedef struct dfft_block_info {
  size_t in_topology[DFFT_MAX_RANK];
  size_t out_topology[DFFT_MAX_RANK];
  size_t in_extents[DFFT_MAX_RANK];
  size_t out_extents[DFFT_MAX_RANK];
  size_t buffer_extents[DFFT_MAX_RANK + 1];
  size_t in_map[DFFT_MAX_RANK];
  size_t out_map[DFFT_MAX_RANK];
  size_t in_axis, out_axis, fft_dim, comm_axis;
  int block_type; /* 0 = Transpose, 1 = FFT */
  size_t block_idx;
} dfft_block_info;

const char* dfft_last_error(void);
int dfft_version(int* major, int* minor);

/* ---- plans ------------------------------------------------------------- */
/* Replaces Plan::Plan + internal_plan_factory (Plan.hpp:46-113, 154-181). */
int dfft_plan_create(const dfft_desc* desc, dfft_plan** plan);
/* Replaces the Plan destructor (buffers, sub-plans, PencilPlan.hpp:1190-1198). */
int dfft_plan_destroy(dfft_plan* plan);
/* Replaces Plan::name() (Plan.hpp:199): "SharedPlan", "SlabPlan", "PencilPlan". */
int dfft_plan_name(const dfft_plan* plan, char* buf, size_t len);
/* Local extents/starts of virtual rank `vrank` for the input (which = 0) or the
 * output (which = 1): compute_local_extents_and_starts (MPI_Extents.hpp:546-577). */
int dfft_plan_local_extents(const dfft_plan* plan, int vrank, int which, size_t* extents, size_t* starts);
/* Block list of virtual rank `vrank` (Slab/PencilBlockAnalysesInternal, BlockAnalyses.hpp:415-506). */
int dfft_plan_num_blocks(const dfft_plan* plan, int* nblocks);
int dfft_plan_block_info(const dfft_plan* plan, int vrank, int idx, dfft_block_info* info);
/* Global maximum of the per-rank buffer sizes in scalar units (m_max_buffer_size). */
int dfft_plan_max_buffer_size(const dfft_plan* plan, size_t* size);
/* Device bytes the plan owns (scratch + twiddle tables). */
int dfft_plan_workspace_bytes(const dfft_plan* plan, size_t* bytes);

/* Replaces execute(plan, in, out, direction) (Plan.hpp:212-238).  Forward
 * reads `in` (P local input arrays) and writes `out` (P local output arrays);
 * backward reads `in` (the P complex spectra) and writes `out` (P arrays of
 * the plan's input shape).  Inputs are not modified. */
int dfft_execute(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream);
/* Adjoint of dfft_execute(direction) under the plain real inner product
 * Re sum conj(a) b of each side (every stored half-spectrum entry counted
 * once): `in` has the shape of that direction's output, `out` of its input.
 * C2C: the opposite transform with this direction's scale; R2C forward:
 * C2R with the interior entries (0 < k < n/2) halved; C2R backward: R2C
 * with them doubled.  No counterpart in the reference -- for gradient-based
 * callers (e.g. the torch.autograd test harness). */
int dfft_execute_adjoint(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream);

/* Number of kernel launches (line passes, remaps) one execute of `direction` issues. */
int dfft_plan_num_passes(const dfft_plan* plan, int direction, int* npasses);
/* dfft_execute that also records a HIP event pair around every launch on
 * `stream`, synchronizes, and writes each launch's duration (ms) into
 * ms_per_pass[0..npasses) -- the per-kernel timing the reference gets from
 * Kokkos::Profiling regions (MPI_All2All.hpp:29-32). */
int dfft_execute_timed(dfft_plan* plan, void* const* in, void* const* out, int direction, void* stream,
                       float* ms_per_pass, int capacity, int* npasses);

/* Launch-shape knobs of a fused plan (no counterpart in the reference, whose
 * vendor FFT plans choose their own):  "chunk" / "fwd.chunk" / "bwd.chunk" =
 * planes per chunk of the first two passes (0 = off); "pad" = rows of the
 * plan scratch padded to 128-byte multiples (-1 auto: when that adds at most
 * 1/16, 0 off, 1 on); "fwd.<i>.lpb",
 * "fwd.<i>.vpt", "fwd.<i>.lf", "fwd.<i>.nt", "fwd.<i>.pair", "fwd.<i>.xcd",
 * "fwd.<i>.swz", "fwd.<i>.split" (and "bwd.…") = lines per block, values
 * per thread (4, 8, 16), lane order, nontemporal global accesses (-1 = on
 * for arrays of at least 1 GiB), two real lines per complex FFT in an R2C /
 * C2R pass (default on), XCD-contiguous block numbering, a column-major walk
 * of the block grid, radix-stage exchanges through LDS in real and
 * imaginary halves (default on; half the LDS per line; FP64 lines and FP32
 * lines of 1024 points or more), and "fwd.<i>.dual" = two adjacent FP32
 * lines per lane with 16-byte accesses (default on; C2C lanes-across-lines
 * passes whose consecutive lines are adjacent row elements; elsewhere
 * ignored), "fwd.<i>.lin" = lines of a one-line C2C lanes-across-lines pass
 * from the padded plan scratch into odd unpadded rows enumerated over the
 * real elements, so blocks write whole aligned segments (default on;
 * elsewhere ignored) of pass i.  A shape that does not fit a workgroup makes
 * execute return DFFT_ERROR_HIP without launching. */
int dfft_plan_set_option(dfft_plan* plan, const char* key, long value);
int dfft_plan_get_option(const dfft_plan* plan, const char* key, long* value);
/* Autotuner: times every launch shape of every pass and the chunk sizes on the
 * given arrays and keeps the fastest.  Like FFTW_MEASURE planning it
 * overwrites the contents of both `in` and `out`. */
int dfft_plan_autotune(dfft_plan* plan, void* const* in, void* const* out, void* stream);

/* ---- exchange layer (standalone, one GPU) ------------------------------ */
/* pack() (PackUnpack.hpp:439-460) of one rank's local array `src` (storage
 * extents src_extents, mapped by src_map) into the exchange buffer along
 * logical `axis`.  `layout` is the views' layout: DFFT_LAYOUT_RIGHT arrays
 * are row-major and their buffer is (nprocs, e_0..e_{rank-1});
 * DFFT_LAYOUT_LEFT arrays are column-major and their buffer is
 * (e_0..e_{rank-1}, nprocs) (PackUnpack.hpp:176-178, Extents.hpp:44-60).
 * buffer_extents (rank + 1 values) are given in that order. */
int dfft_pack(const void* src, const size_t* src_extents, const size_t* src_map, int rank, void* buffer,
              const size_t* buffer_extents, int nprocs, int axis, int layout, int elem_size, void* stream);
/* unpack() (PackUnpack.hpp:478-499): the inverse placement into `dst`. */
int dfft_unpack(const void* buffer, const size_t* buffer_extents, int nprocs, void* dst,
                const size_t* dst_extents, const size_t* dst_map, int rank, int axis, int layout, int elem_size,
                void* stream);
/* all2all() (MPI_All2All.hpp:23-72) among P virtual ranks: block p of rank
 * r's send buffer lands as block r of rank p's receive buffer. */
int dfft_all2all(void* const* send, void* const* recv, int nprocs, size_t block_bytes, void* stream);
/* TransBlock::operator() (TransBlock.hpp:76-107) for all P virtual ranks in
 * one fused remap: the local arrays `in` (distribution `in_topology`, map
 * `in_map`) become `out` (distribution `out_topology`, map `out_map`). */
int dfft_exchange(void* const* in, void* const* out, int rank, const size_t* global_extents,
                  const size_t* in_topology, int in_topology_layout, const size_t* in_map,
                  const size_t* out_topology, int out_topology_layout, const size_t* out_map, int elem_size,
                  void* stream);

/* ---- groups of independent transforms ----------------------------------- */
/* A group runs many single-GPU transforms of possibly different extents
 * (C2C or R2C, one precision, LayoutRight, nranks = 1) together: line pass k
 * of every member is one grouped kernel launch (members whose line length has
 * no grouped kernel get launches of their own).  Replaces a loop of
 * execute(plan_i, in_i, out_i) over separate plans (Plan.hpp:212-238); the
 * reference batches only equal-size fields, through a leading batch axis
 * (navier-stokes-MPI-batched/main.cpp:102-107). */
typedef struct dfft_group dfft_group;
int dfft_group_create(int count, const dfft_desc* descs, dfft_group** group);
int dfft_group_destroy(dfft_group* group);
/* in[i] / out[i]: member i's arrays.  use_graph: the launches are captured
 * once per direction and pointer set into a hipGraph and replayed after. */
int dfft_group_execute(dfft_group* group, void* const* in, void* const* out, int direction, int use_graph,
                       void* stream);
/* dfft_group_execute (use_graph = 0) that allocates no device memory: no
 * counterpart in the reference; for callers that cannot call hipMalloc (code
 * under stream capture, serving loops, harnesses that forbid allocation).
 * The call's pass tables go into a static device array of 4 MiB per GPU
 * (about 2,600 member passes); a call whose tables need more returns
 * DFFT_ERROR_UNSUPPORTED before launching anything.  The array is shared by
 * every group and every call on the device: calls that overlap on one GPU
 * race.  The call synchronises the stream once (after the table copy, before
 * its kernels are queued).  R2C members still use the complex scratch their
 * group allocated at creation. */
int dfft_group_execute_noalloc(dfft_group* group, void* const* in, void* const* out, int direction, void* stream);
int dfft_group_num_launches(const dfft_group* group, int direction, int* nlaunches);
/* "lds_cap" (LDS bytes per block of the grouped kernel, 1024..65536),
 * "grid_mult" (multiples of the resident-block grid; 0 = one block per
 * virtual block), "nt" (nontemporal accesses: -1 auto, 0, 1), "dual"
 * (FP32: two adjacent lines per lane on strided passes where they apply; 1
 * default, 0 off), "lanes" (64, 128 or 256 lanes per block for the
 * launches whose lines are all contiguous; 256 default) and "stage" (FP32:
 * contiguous 16-byte-aligned lines move 16 bytes per lane through LDS; 1
 * default, 0 off) and "contig_nt" (nontemporal accesses of those launches:
 * -1 as "nt", 0, 1); the autotuner times combinations on the caller's arrays
 * (contents are overwritten) and keeps the fastest. */
int dfft_group_set_option(dfft_group* group, const char* key, long value);
int dfft_group_get_option(const dfft_group* group, const char* key, long* value);
int dfft_group_autotune(dfft_group* group, void* const* in, void* const* out, void* stream);

/* ---- pseudo-spectral Navier-Stokes (examples/navier-stokes-MPI) -------- */
/* The example's solver (main.cpp:31-986, utils/math_utils.hpp:47-141) on P =
 * px*py virtual ranks: velocity spectra in the (1,1,px,py) topology of the
 * batched example (navier-stokes-MPI-batched/main.cpp:102-107), fields as a
 * leading axis of 3.  form 0 evaluates u.grad(u) as div(u u) (3 C2R + 6 R2C
 * per RK stage); form 1 is the example's advective form (12 C2R + 3 R2C). */
typedef struct dfft_ns dfft_ns;
typedef struct dfft_ns_desc {
  int n;      /* grid points per direction */
  int px, py; /* virtual process grid */
  double lx;  /* domain [0, lx*pi)^3 */
  double nu;  /* 1/Re */
  double dt;
  int form;   /* 0 divergence form, 1 advective form */
} dfft_ns_desc;
int dfft_ns_create(const dfft_ns_desc* desc, dfft_ns** ns);
int dfft_ns_destroy(dfft_ns* ns);
/* which = 0: physical block (3, ext...) of rank r; 1: spectral block */
int dfft_ns_local_extents(const dfft_ns* ns, int vrank, int which, size_t* extents, size_t* starts);
/* Taylor-Green initial condition (main.cpp:349-353, 398-417) */
int dfft_ns_init_taylor_green(dfft_ns* ns, double v0, void* stream);
/* physical velocity (P blocks of (3, ...) float64) -> dealiased, projected state */
int dfft_ns_set_physical(dfft_ns* ns, void* const* u3, void* stream);
int dfft_ns_set_state(dfft_ns* ns, void* const* y3, void* stream);
int dfft_ns_get_state(dfft_ns* ns, void* const* y3, void* stream);
/* RK4 steps (NavierStokes::solve, main.cpp:544-561); use_graph replays one
 * captured step as a hipGraph after an eager first step */
int dfft_ns_step(dfft_ns* ns, int nsteps, int use_graph, void* stream);
/* time the launch shapes of every plan the solver uses (on its scratch
 * buffers; the state is kept) and keep the fastest -- no counterpart in the
 * reference, whose plans are created once with vendor defaults */
int dfft_ns_autotune(dfft_ns* ns, void* stream);
/* "fuse_products" (1 default, 0): form the divergence form's 6 products in
 * the forward transform's first pass from the velocity lines (no product
 * fields stored) where the plan allows it; "products_fused" (read only): 1
 * when the step runs that way.  No counterpart in the reference. */
int dfft_ns_set_option(dfft_ns* ns, const char* key, long value);
int dfft_ns_get_option(const dfft_ns* ns, const char* key, long* value);
/* kinetic energy, FP64 reduction (diag_energy, main.cpp:855-910) */
int dfft_ns_energy(dfft_ns* ns, double* energy, void* stream);
/* diagnostics (diag_fields / diag_energy, main.cpp:812-935): physical
 * velocity and vorticity of the state into P blocks of (3, ...) float64; the
 * radial spectrum sum 0.5 f |u_k|^2 over num_bins half-open bins of
 * [0, k_max] (f from the global kz index: 1 on kz = 0, 2 elsewhere) with the
 * bin midpoints; and the reference's raw files: per rank r with (rx, ry) =
 * rank_to_coord({px, py}, r), <out_dir>_px<px>_py<py>/<label>_rx###_ry###_
 * ##########.dat for u, v, w, vor_x, vor_y, vor_z (LayoutRight float64
 * blocks), plus E_spec and k_vals (50 bins) under rx000_ry000. */
int dfft_ns_velocity(dfft_ns* ns, void* const* u3, void* stream);
int dfft_ns_vorticity(dfft_ns* ns, void* const* w3, void* stream);
int dfft_ns_energy_spectrum(dfft_ns* ns, int num_bins, double* e_spec, double* k_vals, void* stream);
int dfft_ns_diag(dfft_ns* ns, const char* out_dir, int diag_it, double* energy, void* stream);
/* right-hand side R(y), its tangent R'(y)d and adjoint R'(y)^*l (form 0);
 * no counterpart in the reference -- for gradient-based callers */
int dfft_ns_rhs(dfft_ns* ns, void* const* y3, void* const* out3, void* stream);
int dfft_ns_rhs_tangent(dfft_ns* ns, void* const* y3, void* const* d3, void* const* out3, void* stream);
int dfft_ns_rhs_adjoint(dfft_ns* ns, void* const* y3, void* const* l3, void* const* out3, void* stream);
/* J = kinetic energy after nsteps RK4 steps from the physical velocity u0,
 * and dJ/du0, by checkpointed adjoint recomputation (form 0) */
int dfft_ns_gradient(dfft_ns* ns, int nsteps, int checkpoint_every, void* const* u0, void* const* grad, double* J,
                     void* stream);

/* ---- planner (host only, no device needed) ----------------------------- */
/* get_all_slab_topologies (Topologies.hpp:420-618); topologies written as
 * ntopo x rank entries; *ntopo holds the capacity on entry. */
int dfft_get_all_slab_topologies(int rank, const size_t* in_topology, const size_t* out_topology,
                                 int fft_rank, const int* axes, size_t* topologies, int* ntopo);
/* get_all_pencil_topologies (Topologies.hpp:642-811). */
int dfft_get_all_pencil_topologies(int rank, const size_t* in_topology, int in_layout,
                                   const size_t* out_topology, int out_layout, int fft_rank, const int* axes,
                                   size_t* topologies, size_t* trans_axes, size_t* layouts, int* ntopo);
/* Block analyses of one rank without allocating a plan. */
int dfft_block_analyses(const dfft_desc* desc, int vrank, dfft_block_info* infos, int* nblocks,
                        size_t* max_buffer_size);
/* rank_to_coord (MPI_Extents.hpp:492-529): coordinates of rank `vrank` in a
 * Topology<size_t, rank, topology_layout>; DFFT_ERROR_INVALID_ARGUMENT when
 * vrank >= the topology size. */
int dfft_rank_to_coord(int rank, const size_t* topology, int topology_layout, int vrank, size_t* coord);
/* compute_local_extents_and_starts (MPI_Extents.hpp:546-577): local extents
 * and global starts of rank `vrank`'s block of a global array distributed
 * over the topology (first extent % topology bins one element longer). */
int dfft_compute_local_extents_and_starts(int rank, const size_t* global_extents, const size_t* topology,
                                          int topology_layout, int vrank, size_t* extents, size_t* starts);

#ifdef __cplusplus
}
#endif

#endif /* DFFT_H */
