// Launch interface of the gfx950 kernels of the LGF fast-multipole path.
// Every field is a contiguous array of blocks: field[slot][n^3] FP64, x
// fastest, slots ordered by level then Morton key.  Every spectrum is
// spec[slot][NF] complex FP64 with NF = H*P*P, P = 2n-1, H = n, index
// (kx*P + ky)*P + kz.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iblgf_amd {

// One sibling group of the spectral multiply-accumulate: the (up to) eight
// children of one parent that are targets at this level, and the range of
// source entries shared by them.
struct MacGroup
{
    int tslot[8];  // target slot per child id, -1 if not a target
    int begin;     // first source entry
    int count;     // number of source entries
    int nfull;     // the first nfull entries act on every target of the group
};

// A source entry: block slot, position r = (rx, ry, rz) of the source in the
// 6^3 block region spanned by the children of the parent's 27 neighbours,
// stored as the kernel-row base ob = (rx+1) + 7 (ry+1) + 49 (rz+1) (target
// child c = (cx, cy, cz) uses offset row ob - (cx + 7 cy + 49 cz)), and the
// 8-bit mask of the group's targets that it acts on.
struct MacSrc
{
    int slot;
    int meta;  // ob | (tmask << 9)
};

// k components of one tree in one launch (grid dimension y for the block
// kernels, a component loop for the multiply-accumulate): component c of a
// field array starts c*fstride doubles in, of a spectrum array c*sstride
// complex values in.  The default is one component.
struct Batch
{
    int  k = 1;
    long fstride = 0;
    long sstride = 0;
    // Launch shapes chosen by the solver's autotuner (results do not depend
    // on them): mac_policy 0 spreads a full-size multiply-accumulate launch
    // over the resident workgroups in rounds, 1 gives every frequency chunk a
    // workgroup, 2 runs the resident workgroups as a persistent grid that
    // claims (chunk, group run) units from a device work queue; dft_grid 0
    // runs the block transforms on the resident
    // workgroups looping over blocks, 1 gives every (block, component) a
    // workgroup.
    int mac_policy = 0;
    int dft_grid = 0;
    // Field index of component 0 (vector fields: component c is field index
    // (comp0 + c) % 3), for the mesh-type level transfers.
    int comp0 = 0;
};

// Mesh types of the level transfers (MeshObject of the reference).
constexpr int kMeshCell = 0, kMeshFace = 1, kMeshEdge = 2;

// Per-component field base pointers, passed by value (component c in p[c]).
// Bit c of f32 set: component c is an FP32 field and p[c] holds its float
// address (read and written as float, converted to and from FP64 by the
// copies between user fields and a solve's work arrays).
constexpr int kMaxFieldTable = 64;
struct FieldTable
{
    double*            p[kMaxFieldTable];
    unsigned long long f32 = 0;
};

// Bytes per staged kernel row of the multiply-accumulate (KF frequencies of
// one offset: 8 FP64 or 16 FP32 complex values), and the largest offset count
// of a compact launch.
constexpr int kMacRowBytes = 128;
constexpr int kMacCompactMax = 200;

// Arithmetic of the block transforms (the solver's precision codes): 0 FP64
// (VALU), 1 FP64 legs with FP32 spectra (VALU), and the matrix-core legs with
// FP32 spectra: BF16 three-term split, FP16 three-term split with power-of-two
// line scaling, FP16 one term with line scaling, four signed 7-bit integer
// digits with exact INT32 accumulation.
constexpr int kPrecFp64 = 0, kPrecMixed = 1, kPrecBf16x3 = 2, kPrecF16x3 = 3, kPrecF16 = 4, kPrecI8x4 = 5;
// Solver precision code: FP16 three-term forward legs, FP64 inverse legs.
constexpr int kPrecF16x3Forward = 6;
constexpr bool prec_mfma(int p) { return p >= kPrecBf16x3 && p <= kPrecI8x4; }

namespace launch {

bool supported_n(int n);

void zero_blocks(hipStream_t s, int n, double* f, const int* slots, int count, const Batch& b = {});
// Whole blocks; interior_only copies the interior points and writes zeros on
// the one-point boundary ring of dst.
void copy_blocks(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count,
    bool interior_only, const Batch& b = {});
// The same for k components whose fields sit at unrelated addresses, in one launch.
void copy_blocks_table(hipStream_t s, int n, const FieldTable& dst, const FieldTable& src, const int* slots,
    int count, bool interior_only, int k);
void axpy_blocks(hipStream_t s, int n, double* dst, const double* src, const int* slots, int count,
    double a, const Batch& b = {});

// dst[entry.parent] (=|+=) sum_c K(child) for entries {parent, child[8]}; K
// built from side matrices (row = output point, col = input point):
// mesh < 0, mats[2][n*n] by the child's bit on each axis (the FMM's Lagrange
// operators); mesh = kMeshCell / kMeshFace / kMeshEdge, mats[2 relative
// positions][2][n*n] with each axis's position set by the mesh type and the
// component's field index (b.comp0 + component) % 3.
// children_interior: the children's boundary rings are read as zero.
void transfer_gather(hipStream_t s, int n, const double* src, double* dst, const int* entries, int count,
    const double* mats, bool accumulate, const Batch& b = {}, bool children_interior = false, int mesh = -1);
// dst[entry.child] (=|+=) K(src[entry.parent]) for entries {child, parent, cidx}.
void transfer_scatter(hipStream_t s, int n, const double* src, double* dst, const int* entries, int count,
    const double* mats, bool accumulate, const Batch& b = {}, int mesh = -1);

void buffer_exchange(hipStream_t s, int n, double* f, const int* slots, int count, const int* neighbor,
    const Batch& b = {});
void correction_stencil(hipStream_t s, int n, const double* target_tmp, double* corr, const int* slots,
    int count, double inv_dx2, double omega2, const Batch& b = {});

// prec 0: every spectrum FP64 (double2);
// prec 1: source, kernel and target spectra stored FP32 (float2); the MAC
// forms FP32 products, accumulates them in FP64 and rounds each target
// spectrum once on store; DFT arithmetic, LDS intermediates and fields FP64;
// prec 2-5 (prec_mfma): spectra and MAC as prec 1, the DFT legs on the
// matrix cores (forward_dft_mfma / inverse_dft_mfma).
// Blocks slots[split..count) are read from field_in with their boundary ring
// taken as zero, the others from field.
void forward_dft(hipStream_t s, int n, int prec, const double* field, const int* slots, int count, void* spec,
    const Batch& b = {}, const double* field_in = nullptr, int split = -1);
// Writes (accumulate: adds into) every point of each listed field block.
void inverse_dft(hipStream_t s, int n, int prec, const void* spec, const int* slots, int count, double* field,
    double scale, bool accumulate, const Batch& b = {});
// olist[nused]: the offsets (0..342, o = (ox+3) + 7 (oy+3) + 49 (oz+3)) the
// launch's entries use.  A compact launch (nused <= kMacCompactMax) stages
// only those rows, row i = olist[i]; otherwise row o = offset o.  Row nrows
// (nused or 343) is all zeros.  rowoff[e] holds entry e's 8 LDS byte offsets
// row * kMacRowBytes (target child t in 16-bit field t; children outside the
// entry's mask point at the zero row).
// queue: the launch's nine work-queue counters (zero between launches),
// used when b.mac_policy is 2.
void spectral_mac(hipStream_t s, int n, int prec, const MacGroup* groups, int ngroups, const MacSrc* srcs,
    const uint4* rowoff, const int* olist, int nused, bool compact, const void* ghat, const void* sspec,
    void* tspec, const Batch& b = {}, int* queue = nullptr);
void to_float2(hipStream_t s, const double2* in, float2* out, long n);

// The forward and inverse transforms with every 1-D leg on the matrix cores
// (FP32 spectra, FP32 intermediates); prec is one of the split precisions
// below, the arguments are those of forward_dft / inverse_dft.
void forward_dft_mfma(hipStream_t s, int n, int prec, const double* field, const int* slots, int count,
    float2* spec, const Batch& b = {}, const double* field_in = nullptr, int split = -1);
void inverse_dft_mfma(hipStream_t s, int n, int prec, const float2* spec, const int* slots, int count,
    double* field, double scale, bool accumulate, const Batch& b = {});

// Kernel spectra for the 343 offsets o in [-3,3]^3 at level difference d:
// block[a] = K(step*(-o*e - (n-1) + a)), a in [0,P)^3, step = 2^d, then r2c.
// kind 0: Laplace LGF (table g up to nmax, asymptotics beyond);
// kind 1: integrating factor, iv[k] = e^{-2a} I_k(2a) for k <= kmax.
void lgf_spectra(hipStream_t s, int n, int kind, int d, const double* g, int nmax, const double* iv,
    int kmax, double* sample_scratch, double2* scratch_a, double2* scratch_b, double2* ghat);

}  // namespace launch
}  // namespace iblgf_amd
