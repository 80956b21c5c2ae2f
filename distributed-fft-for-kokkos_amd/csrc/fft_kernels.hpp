// Host-visible launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace dfft {

enum LineKind : int32_t { kC2C = 0, kR2C = 1, kC2R = 2 };

// One batched 1-D transform pass along logical axis `axis`: every line of the
// index space (all axes but `axis`) is read through descriptor `in`,
// transformed and written through `out`.  The two descriptors may describe
// different distributions (that is how the exchange between topologies is
// fused into the first and last pass of a plan).
struct LinePass {
  DDesc in, out;
  int32_t axis;
  int32_t nbatch;                       // ndim - 1
  int32_t batch_axis[kMaxDim - 1];      // slowest ... fastest
  int32_t batch_ext[kMaxDim - 1];
  FastDiv batch_div[kMaxDim - 1];
  int32_t nlines;                       // lines processed by this launch
  int32_t line_begin;                   // first line (index space of all lines)
  int32_t kind;                         // LineKind
  int32_t n;                            // transform length (real length for R2C/C2R)
  int32_t inverse;                      // 1: exponent +2 pi i jk/n
  int32_t lines_fast;                   // 1: consecutive lanes walk across lines
  int32_t lines_per_block;              // 0: launcher heuristic
  int32_t values_per_thread;            // 0: default (8); compile-time kernels exist for 4, 8, 16
  const void* tw;                       // exp(-2 pi i e / n), e in [0, n), Real precision
  double scale;
  // 1: an R2C (forward) or C2R (inverse) pass on the fast kernel transforms
  // two real lines per complex FFT (lines 2q and 2q+1 of the pass)
  int32_t paired;
  // pad_real > 0: batch axis pad_axis is enumerated over a padded extent
  // (batch_ext, the row pitch of a padded array) and lines whose index along
  // it is >= pad_real are skipped, so blocks of lines start on aligned rows
  int32_t pad_axis;
  int32_t pad_real;
  // 1: consecutive blocks step along the slower batch axes first (the block
  // grid, rows of blocks along the fastest batch axis, is walked column by
  // column), spreading concurrent blocks over more of the array
  int32_t swizzle;
  // 1: blocks are renumbered so each of the 8 XCDs (which take workgroups
  // round-robin) walks one contiguous range of lines
  int32_t xcd_remap;
  // 1: global loads and stores of the line data are nontemporal (streaming
  // arrays larger than the caches; measured 5-13 % faster per pass at 512^3)
  int32_t nontemporal;
  // 1 (FP64 fast kernels, FP32 ones of length >= 1024): the radix stages
  // exchange real parts, then imaginary parts, through a line buffer of
  // reals, halving the LDS of a line (more resident blocks per CU for a few
  // more barriers)
  int32_t split_lds;
  // 1 (FP32 C2C lanes-across-lines passes whose consecutive lines are
  // adjacent elements of one row): two lines per lane, one 16-byte access
  // per element pair; ignored where the lines are not adjacent
  int32_t dual_lines;
  // two-lines-per-lane passes only: bit 0 set when the input side's line
  // pairs are loaded as two 8-byte elements instead of one aligned 16-byte
  // pair -- an input with odd rows under a padded enumeration, whose pair at
  // a row's end has its second line in the padding (that element is then not
  // loaded)
  int32_t dual_split;
  // 1 (one-line lanes-across-lines C2C passes from a padded scratch into odd
  // unpadded output rows): lines enumerated over the real elements only, so a
  // block writes whole, aligned output segments (see lines_linear_unpadded)
  int32_t linear_lines;
  // FP64 R2C passes only (pair kernel): 1 = the line of field f (the batch
  // axis prod_axis, six fields) is the pointwise product u_I * u_J of two of
  // the three fields that the input descriptor holds along that axis, (I, J)
  // = (0,0) (0,1) (0,2) (1,1) (1,2) (2,2) for f = 0..5 -- the products of the
  // divergence-form Navier-Stokes right-hand side formed as the lines load
  int32_t products;
  int32_t prod_axis;
  // Weight of the interior half-spectrum entries (0 < k < n/2) of an R2C /
  // C2R line, applied to the values a C2R line loads or an R2C line stores;
  // 0 means 1.  Adjoint transforms use 1/2 (C2R) and 2 (R2C).
  double herm_weight;
  // long_n1 > 0: lines too long for one LDS buffer, n = n1 * n2, run as two
  // launches through long_scratch (complex, nlines * n values): first the n2
  // sub-lines of stride n2 and length n1 of every line, each output k1
  // multiplied by exp(-2 pi i i2 k1 / n); then the n1 contiguous sub-lines of
  // length n2, whose output k2 is element k1 + n1 k2 of the line
  int32_t long_n1;
  const void* long_tw1;  // exp(-2 pi i e / n1)
  const void* long_tw2;  // exp(-2 pi i e / n2)
  void* long_scratch;
};

// Element-wise permuted copy between two distributions of the same logical
// array (one-GPU form of pack -> all-to-all -> unpack, and of local
// transposes).  Iterates the destination in its storage order.
struct RemapArgs {
  DDesc src, dst;
  int64_t total;
  int32_t iter_axis[kMaxDim];  // slowest ... fastest
  int32_t iter_ext[kMaxDim];
  FastDiv iter_div[kMaxDim];
};

// Grouped line pass (one launch for the same pass of many transforms, line
// lengths 16, 32, 64, 96, 128, 160, 256 in any mix).  kGroupThreads lanes per
// block; launches of contiguous lines (lanes along the line) may use fewer
// (64 or 128: smaller blocks, cheaper barriers, more blocks per CU).
constexpr int kGroupThreads = 256;
constexpr size_t kGenericLdsMax = 160 * 1024;  // LDS of one gfx950 workgroup
constexpr size_t kGroupLds = 64 * 1024;  // dynamic LDS of a grouped launch, at most
// lines per block and dynamic LDS bytes of the grouped kernel for length n
// in blocks of `lanes` lanes (64, 128 or kGroupThreads) with at most lds_cap
// bytes of LDS per block; false if n has no grouped kernel
bool group_line_shape(int n, bool fp64, size_t lds_cap, int lanes, int* lines_per_block, size_t* lds_bytes);
// the same for FP32 lines taken two per lane (false if n has no such kernel)
bool group_dual_shape(int n, size_t lds_cap, int* lines_per_block, size_t* lds_bytes);
// true if two adjacent lines per lane apply to pass p with that many lines per
// block (FP32 C2C, lanes across lines, consecutive lines adjacent row
// elements on both sides, 16-byte aligned rank buffers; grouped passes)
bool dual_lines_apply(const LinePass& p, int lines_per_block);
// grid of the persistent grouped launch: resident blocks of the whole device
int group_grid(size_t lds_bytes, bool fp64, bool dual, int lanes);
// dual: the launch holds only members with dual_lines set (FP32, kGroupThreads
// lanes); lanes: lanes per block of a one-line launch; stage (FP32 one-line):
// contiguous aligned lines move 16 bytes per lane through LDS
hipError_t launch_group_pass(const LinePass* items, const int32_t* first_block, int count, int total_blocks,
                             int grid, size_t lds_bytes, bool fp64, bool nontemporal, bool dual, int lanes,
                             bool stage, hipStream_t s);
bool has_fast_line_kernel(int n);
// true if a line of length n runs as one pass: the fast set, or the generic
// kernel (one line of n complex values, twice, in at most 160 KB of LDS)
bool line_length_single_pass(int n, bool fp64);
// n1 of a two-pass split n = n1 * n2 with both factors single-pass lengths
// (the divisor nearest sqrt(n), n1 <= n2), 0 if n has none
int long_line_split(int n, bool fp64);
// single pass or a two-pass split
bool line_length_supported(int n, bool fp64);
// true where the plan runs length n by Bluestein's algorithm: no other kernel
// runs it, or a measured cost model puts Bluestein well below the generic
// kernel (large prime factors)
bool bluestein_preferred(int n, bool fp64);
hipError_t launch_line_pass(const LinePass& p, bool fp64, hipStream_t s);
// Bluestein steps of pass p over an (nlines, 2^log2m) complex scratch (see
// fft_kernels.hip); chirp: exp(-pi i e / n), e in [0, 2n), FP64; filt: the
// m-point transform of the chirp filter divided by m, FP64
hipError_t launch_bluestein_pre(const LinePass& p, int log2m, const void* chirp, void* s1, bool fp64, hipStream_t s);
hipError_t launch_bluestein_mul(int64_t total, int log2m, const void* filt, void* buf, bool fp64, hipStream_t s);
hipError_t launch_bluestein_post(const LinePass& p, int log2m, const void* chirp, const void* s1, bool fp64,
                                 hipStream_t s);
// an FP64 R2C pass with p.products set (paired lines, lengths 8..1024 powers
// of two); hipErrorNotSupported for other passes
hipError_t launch_products_pass(const LinePass& p, hipStream_t s);
bool products_pass_supported(int n);
hipError_t launch_remap(const RemapArgs& a, int esize, hipStream_t s);

// Reference-exact pack/unpack of one rank's local array (mapped by `map`)
// into/out of a (P, e_0..e_{D-1}) LayoutRight buffer, or, with layout_left,
// a LayoutLeft array into/out of an (e_0..e_{D-1}, P) LayoutLeft buffer.
struct PackArgs {
  int32_t ndim;
  int32_t layout_left;
  int32_t nprocs;
  int32_t axis;           // logical axis split (pack) / merged (unpack)
  int32_t merged_size;    // local extent of that axis
  int32_t map[kMaxDim];
  int64_t local_ext[kMaxDim];   // storage extents of the local array
  int64_t buf_ext[kMaxDim + 1]; // (P, e_0..e_{D-1}), or (e_0..e_{D-1}, P) with layout_left
  int64_t total;                // elements of the buffer
  const char* src;
  char* dst;
};
hipError_t launch_pack(const PackArgs& a, int esize, bool unpack, hipStream_t s);

// Device-local all-to-all of P buffers of shape (P, block): recv[q][p] = send[p][q].
struct All2AllArgs {
  int32_t nprocs;
  int64_t block_bytes;
  const char* send[kMaxRanks];
  char* recv[kMaxRanks];
};
hipError_t launch_all2all(const All2AllArgs& a, hipStream_t s);

}  // namespace dfft
