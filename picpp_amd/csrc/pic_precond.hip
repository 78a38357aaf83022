// Learned preconditioner: a 3-D convolution U-Net (3 levels, 16/32/64 channels, 3x3x3
// kernels, periodic) in bfloat16 with float32 accumulation on MFMA.
//
// Layout: activations are channels-last bf16 [voxel][channel], voxel = (x*d1 + y)*d2 + z,
// the same order as the solver's owned-node vectors [3][V]. A 3x3x3 convolution is an
// implicit GEMM: rows are voxels, columns output channels, and the reduction runs over
// (tap, input channel) in chunks of 32 fed to v_mfma_f32_16x16x32_bf16. A lane's 8 A
// values are 8 consecutive channels of one neighbour voxel (one 16-byte load), its 8 B
// values 8 consecutive input channels of one output channel (weights stored W[t][co][ci]).
// The input gradient of a convolution is the same kernel on the adjoint weights
// W'[t][ci][co] = W[26 - t][co][ci] (tap 26 - t is offset -o_t). Weight gradients are
// float32 outer-product sums over voxel chunks, reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "pic_precond.hpp"
#include "pic_solver.hpp"

namespace picgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kHeadIn = 16, kHeadOut = 3;
constexpr int kRedBlocks = 512;

#define PC_CHECK(expr)                                                                                    \
	do {                                                                                                  \
		hipError_t e_ = (expr);                                                                           \
		if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
	} while (0)

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even
__device__ inline uint16_t f2bf(float f) {
	uint32_t u = __float_as_uint(f);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

__device__ inline int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline bf16x8 load8(const uint16_t* p) {
	uint4 q = *reinterpret_cast<const uint4*>(p);
	return __builtin_bit_cast(bf16x8, q);
}

// 8 bf16 values from 16 bytes and back
__device__ inline void unpack8(const uint4 q, float f[8]) {
	const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		f[2 * i] = __uint_as_float(w[i] << 16);
		f[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
	}
}
__device__ inline uint4 pack8(const float f[8]) {
	uint32_t w[4];
#pragma unroll
	for (int i = 0; i < 4; ++i) w[i] = f2bf(f[2 * i]) | (uint32_t)f2bf(f[2 * i + 1]) << 16;
	return make_uint4(w[0], w[1], w[2], w[3]);
}

// ------------------------------------------------------------------ convolution (MFMA)
// Y[v][co] = sum_t sum_ci X[v + o_t][ci] W[t][co][ci] (+ b[co], relu); 32 voxels per wave.
template <int CIN, int COUT, class OutT>
__global__ void __launch_bounds__(256) k_conv3(const uint16_t* __restrict__ X, const uint16_t* __restrict__ W,
                                               const float* __restrict__ bias, int relu, OutT* __restrict__ Y, int d0, int d1,
                                               int d2) {
	constexpr int NT = COUT / 16;
	constexpr int NCH = (27 * CIN + 31) / 32;
	const int V = d0 * d1 * d2;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int vbase = (blockIdx.x * 4 + wave) * 32;
	const int r16 = lane & 15, kg = lane >> 4;
	int vx[2], vy[2], vz[2];
#pragma unroll
	for (int h = 0; h < 2; ++h) {
		int v = vbase + 16 * h + r16;
		v = v < V ? v : 0;
		vz[h] = v % d2;
		vy[h] = (v / d2) % d1;
		vx[h] = v / (d1 * d2);
	}
	f32x4 acc[2][NT];
#pragma unroll
	for (int h = 0; h < 2; ++h)
#pragma unroll
		for (int n = 0; n < NT; ++n) acc[h][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	for (int c = 0; c < NCH; ++c) {
		const int kf = c * 32 + kg * 8;
		const int t = kf / CIN, ci = kf % CIN;
		bf16x8 a[2], b[NT];
		if (t < 27) {
			const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				int idx = (wrap(vx[h] + dx, d0) * d1 + wrap(vy[h] + dy, d1)) * d2 + wrap(vz[h] + dz, d2);
				a[h] = load8(X + (long long)idx * CIN + ci);
			}
#pragma unroll
			for (int n = 0; n < NT; ++n) b[n] = load8(W + ((long long)t * COUT + n * 16 + r16) * CIN + ci);
		} else {
#pragma unroll
			for (int h = 0; h < 2; ++h) a[h] = bf16x8{};
#pragma unroll
			for (int n = 0; n < NT; ++n) b[n] = bf16x8{};
		}
#pragma unroll
		for (int h = 0; h < 2; ++h)
#pragma unroll
			for (int n = 0; n < NT; ++n) acc[h][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[h], b[n], acc[h][n], 0, 0, 0);
	}
	// D layout: lane (kg, r16) holds rows 4 kg + i, column r16
#pragma unroll
	for (int h = 0; h < 2; ++h)
#pragma unroll
		for (int n = 0; n < NT; ++n)
#pragma unroll
			for (int i = 0; i < 4; ++i) {
				const int v = vbase + 16 * h + 4 * kg + i;
				const int co = n * 16 + r16;
				if (v >= V) continue;
				float val = acc[h][n][i] + (bias ? bias[co] : 0.0f);
				if (relu) val = val > 0.0f ? val : 0.0f;
				if constexpr (sizeof(OutT) == 2)
					Y[(long long)v * COUT + co] = f2bf(val);
				else
					Y[(long long)v * COUT + co] = val;
			}
}

// The same convolution on bricks staged in LDS. A workgroup of BX x NG waves computes a brick
// of BX x BY x 16 output voxels: wave (w, g) takes x = x0 + w and the BY rows y = y0 + r, each
// row the 16 voxels z0 .. z0 + 15 along z, for the g-th NG-th of the output channels. The input box around the brick, (BX + 2) x (BY + 2) x
// 18 voxels with periodic wrap, is loaded once into LDS (16-byte pieces of 8 channels),
// so the 27 taps read LDS at offsets that are constants of the unrolled chunk loop, and each
// chunk's weights (one 16-byte load per lane, shared by the BY rows) come from L1. The MFMA
// takes the weights as its first operand, so a lane of the result holds 4 consecutive output
// channels of one voxel and stores them at once. Same sums in the same order as k_conv3.
// Needs d0 % BX = d1 % BY = d2 % 16 = 0.
template <int CIN, int COUT, int BX, int BY, int NG, class OutT>
__global__ void __launch_bounds__(64 * BX * NG) k_conv3_brick(const uint16_t* __restrict__ X, const uint16_t* __restrict__ W,
                                                         const float* __restrict__ bias, int relu, OutT* __restrict__ Y, int d0, int d1,
                                                         int d2) {
	constexpr int HX = BX + 2, HY = BY + 2, HZ = 18, CV = CIN / 8;
	constexpr int NT = COUT / 16 / NG;  // output-channel tiles per wave
	constexpr int NCH = (27 * CIN + 31) / 32;
	constexpr int NVOX = HX * HY * HZ;
	__shared__ uint4 box[NVOX * CV];
	const int nbz = d2 / 16, nby = d1 / BY;
	int b = blockIdx.x;
	const int z0 = (b % nbz) * 16;
	b /= nbz;
	const int y0 = (b % nby) * BY, x0 = (b / nby) * BX;
	// one voxel (CV pieces) per thread and pass; in LDS each row (hx, hy) holds its HZ voxels'
	// piece 0, then their piece 1, ...: the 16 lanes of a lane group read 16 consecutive
	// voxels' pieces, 256 contiguous bytes (no bank conflicts)
	for (int vox = threadIdx.x; vox < NVOX; vox += 64 * BX * NG) {
		const int hz = vox % HZ, hxy = vox / HZ;
		const int hy = hxy % HY, hx = hxy / HY;
		const int gx = wrap(x0 - 1 + hx, d0), gy = wrap(y0 - 1 + hy, d1), gz = wrap(z0 - 1 + hz, d2);
		const uint4* src = reinterpret_cast<const uint4*>(X + (((long long)gx * d1 + gy) * d2 + gz) * CIN);
#pragma unroll
		for (int q = 0; q < CV; ++q) box[(hxy * CV + q) * HZ + hz] = src[q];
	}
	__syncthreads();
	const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) % BX, ng = (threadIdx.x >> 6) / BX;
	const int r16 = lane & 15, kg = lane >> 4;
	// the lane's voxel of tap (0, 0, 0) of row 0, piece 0; chunk c adds the tap's and the
	// piece's constant offsets
	const int lane_base = ((w + 1) * HY + 1) * CV * HZ + 1 + r16;
	f32x4 acc[BY][NT];
#pragma unroll
	for (int r = 0; r < BY; ++r)
#pragma unroll
		for (int n = 0; n < NT; ++n) acc[r][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	// chunk c's operands: A pieces of the BY rows from LDS, the weights from L1
	auto load_chunk = [&](const int c, bf16x8 (&a)[BY], bf16x8 (&bw)[NT]) {
		// kf = 32 c + 8 kg = t CIN + ci: for CIN = 16 the two halves of the wave take taps 2c
		// and 2c + 1; otherwise the chunk holds one tap
		const int h = CIN == 16 ? kg >> 1 : 0;
		const int t = CIN == 16 ? 2 * c + h : (c * 32) / CIN;
		const int ci = CIN == 16 ? (kg & 1) * 8 : (c * 32) % CIN + kg * 8;
		const int t0 = CIN == 16 ? 2 * c : t, t1 = CIN == 16 ? 2 * c + 1 : t;
		constexpr int sy = CV * HZ, sx = HY * sy;
		const int o0 = (t0 / 9 - 1) * sx + ((t0 / 3) % 3 - 1) * sy + (t0 % 3 - 1);
		const int o1 = (t1 / 9 - 1) * sx + ((t1 / 3) % 3 - 1) * sy + (t1 % 3 - 1);
		const int off = (h ? o1 : o0) + (ci / 8) * HZ;
		const bool valid = t < 27;
#pragma unroll
		for (int r = 0; r < BY; ++r)
			a[r] = valid ? __builtin_bit_cast(bf16x8, box[lane_base + off + r * sy]) : bf16x8{};
#pragma unroll
		for (int n = 0; n < NT; ++n)
			bw[n] = valid ? load8(W + ((long long)(valid ? t : 0) * COUT + (ng * NT + n) * 16 + r16) * CIN + ci) : bf16x8{};
	};
	// double-buffered: chunk c + 1's loads are issued before chunk c's MFMAs
	bf16x8 a[2][BY], bw[2][NT];
	load_chunk(0, a[0], bw[0]);
#pragma unroll
	for (int c = 0; c < NCH; ++c) {
		const int cur = c & 1;
		if (c + 1 < NCH) load_chunk(c + 1, a[cur ^ 1], bw[cur ^ 1]);
#pragma unroll
		for (int r = 0; r < BY; ++r)
#pragma unroll
			for (int n = 0; n < NT; ++n)
				acc[r][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[cur][n], a[cur][r], acc[r][n], 0, 0, 0);
	}
	// result layout: lane (kg, r16) holds output channels n 16 + 4 kg + i of voxel z0 + r16
	float bv[NT][4];
#pragma unroll
	for (int n = 0; n < NT; ++n)
#pragma unroll
		for (int i = 0; i < 4; ++i) bv[n][i] = bias ? bias[(ng * NT + n) * 16 + 4 * kg + i] : 0.0f;
	OutT* y0p = Y + ((((long long)(x0 + w) * d1 + y0) * d2 + z0 + r16) * COUT + 4 * kg);
#pragma unroll
	for (int r = 0; r < BY; ++r) {
		OutT* yp = y0p + (long long)r * d2 * COUT;
#pragma unroll
		for (int n = 0; n < NT; ++n) {
			float val[4];
#pragma unroll
			for (int i = 0; i < 4; ++i) {
				val[i] = acc[r][n][i] + bv[n][i];
				if (relu) val[i] = val[i] > 0.0f ? val[i] : 0.0f;
			}
			if constexpr (sizeof(OutT) == 2) {
				uint2 pk = make_uint2(f2bf(val[0]) | (uint32_t)f2bf(val[1]) << 16, f2bf(val[2]) | (uint32_t)f2bf(val[3]) << 16);
				*reinterpret_cast<uint2*>(yp + (ng * NT + n) * 16) = pk;
			} else {
				*reinterpret_cast<float4*>(yp + (ng * NT + n) * 16) = make_float4(val[0], val[1], val[2], val[3]);
			}
		}
	}
}

// Brick shapes: BX x NG waves, the NG wave groups splitting the output channels. 8 waves of 4
// rows for 16 -> 16 channels (measured at 128^3: 44.8 us against 52.2 us with 4 waves and
// 54.5 us with 4 waves of 8 rows), 4 waves of 4 rows for the middle level, 2 x 2 bricks for
// the 64-channel layers (the coarsest level: 32^3 at 128^3), with the output tiles spread
// over the waves so that the few bricks still give every SIMD several waves.
template <int CIN, int COUT>
struct BrickShape {
	static constexpr bool coarse = CIN == 64 || COUT == 64;
	static constexpr int BX = coarse ? 2 : (CIN == 16 && COUT == 16 ? 8 : 4);
	static constexpr int BY = coarse ? 2 : 4;
	static constexpr int NG = coarse ? COUT / 16 : 1;
};

template <int CIN, int COUT, class OutT>
void conv3(const uint16_t* X, const uint16_t* W, const float* b, bool relu, OutT* Y, const int d[3], hipStream_t s) {
	long long V = (long long)d[0] * d[1] * d[2];
	constexpr int BX = BrickShape<CIN, COUT>::BX, BY = BrickShape<CIN, COUT>::BY, NG = BrickShape<CIN, COUT>::NG;
	if (d[0] % BX == 0 && d[1] % BY == 0 && d[2] % 16 == 0) {
		unsigned blocks = (unsigned)(V / (BX * BY * 16));
		hipLaunchKernelGGL((k_conv3_brick<CIN, COUT, BX, BY, NG, OutT>), dim3(blocks), dim3(64 * BX * NG), 0, s, X, W, b, relu ? 1 : 0,
		                   Y, d[0], d[1], d[2]);
		return;
	}
	unsigned blocks = (unsigned)((V + 127) / 128);
	hipLaunchKernelGGL((k_conv3<CIN, COUT, OutT>), dim3(blocks), dim3(256), 0, s, X, W, b, relu ? 1 : 0, Y, d[0], d[1], d[2]);
}

// dispatch on the channel pair
template <class OutT>
void conv3_any(int cin, int cout, const uint16_t* X, const uint16_t* W, const float* b, bool relu, OutT* Y, const int d[3],
               hipStream_t s) {
	if (cin == 16 && cout == 16) return conv3<16, 16, OutT>(X, W, b, relu, Y, d, s);
	if (cin == 16 && cout == 32) return conv3<16, 32, OutT>(X, W, b, relu, Y, d, s);
	if (cin == 32 && cout == 64) return conv3<32, 64, OutT>(X, W, b, relu, Y, d, s);
	if (cin == 64 && cout == 32) return conv3<64, 32, OutT>(X, W, b, relu, Y, d, s);
	if (cin == 32 && cout == 16) return conv3<32, 16, OutT>(X, W, b, relu, Y, d, s);
	throw std::runtime_error("conv3: unsupported channel pair");
}

// ------------------------------------------------------------------ weight gradient
// partial[chunk][t][co][ci] = sum over the chunk's voxels of dZ[v][co] X[v + o_t][ci]; the
// blocks with t = 27 write the bias sums partial[chunk][27 COUT CIN + co].
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(const uint16_t* __restrict__ dZ, const uint16_t* __restrict__ X, int d0, int d1,
                                               int d2, int chunk, float* __restrict__ partial) {
	constexpr int NP = COUT * CIN;
	constexpr int PER = (NP + 255) / 256;
	constexpr int STRIDE = 27 * NP + COUT;
	__shared__ float sz[64][COUT];
	__shared__ float sx[64][CIN];
	const int V = d0 * d1 * d2;
	const int t = blockIdx.x;
	const int v0 = blockIdx.y * chunk;
	const int v1 = v0 + chunk < V ? v0 + chunk : V;
	const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
	float acc[PER];
#pragma unroll
	for (int q = 0; q < PER; ++q) acc[q] = 0.0f;
	for (int vb = v0; vb < v1; vb += 64) {
		for (int e = threadIdx.x; e < 64 * COUT; e += 256) {
			int v = vb + e / COUT;
			sz[e / COUT][e % COUT] = v < v1 ? bf2f(dZ[(long long)v * COUT + e % COUT]) : 0.0f;
		}
		if (t < 27)
			for (int e = threadIdx.x; e < 64 * CIN; e += 256) {
				int v = vb + e / CIN;
				float val = 0.0f;
				if (v < v1) {
					int z = v % d2, y = (v / d2) % d1, x = v / (d1 * d2);
					int idx = (wrap(x + dx, d0) * d1 + wrap(y + dy, d1)) * d2 + wrap(z + dz, d2);
					val = bf2f(X[(long long)idx * CIN + e % CIN]);
				}
				sx[e / CIN][e % CIN] = val;
			}
		__syncthreads();
		if (t < 27) {
			for (int j = 0; j < 64; ++j)
#pragma unroll
				for (int q = 0; q < PER; ++q) {
					int p = threadIdx.x + 256 * q;
					if (p < NP) acc[q] += sz[j][p / CIN] * sx[j][p % CIN];
				}
		} else if (threadIdx.x < COUT) {
			for (int j = 0; j < 64; ++j) acc[0] += sz[j][threadIdx.x];
		}
		__syncthreads();
	}
	float* out = partial + (long long)blockIdx.y * STRIDE;
	if (t < 27) {
#pragma unroll
		for (int q = 0; q < PER; ++q) {
			int p = threadIdx.x + 256 * q;
			if (p < NP) out[(long long)t * NP + p] = acc[q];
		}
	} else if (threadIdx.x < COUT) {
		out[27 * NP + threadIdx.x] = acc[0];
	}
}

// Channel-major copy Xc[c][v] = X[v][c] (bf16), 64 voxels per workgroup through LDS; V % 64 = 0.
template <int C>
__global__ void __launch_bounds__(256) k_channel_major(const uint16_t* __restrict__ X, long long V, uint16_t* __restrict__ Xc) {
	constexpr int S = C + 2;  // row stride of the LDS tile (2-byte words)
	__shared__ uint32_t tile[64 * S / 2];
	uint16_t* t16 = reinterpret_cast<uint16_t*>(tile);
	const long long v0 = (long long)blockIdx.x * 64;
	for (int e = threadIdx.x; e < 64 * C / 8; e += 256) {
		const int vl = e / (C / 8), q = e % (C / 8);
		const uint4 d = reinterpret_cast<const uint4*>(X + (v0 + vl) * C)[q];
		uint32_t* row = tile + (vl * S + q * 8) / 2;
		row[0] = d.x;
		row[1] = d.y;
		row[2] = d.z;
		row[3] = d.w;
	}
	__syncthreads();
	for (int e = threadIdx.x; e < C * 8; e += 256) {
		const int c = e / 8, seg = e % 8;
		uint32_t w[4];
#pragma unroll
		for (int j = 0; j < 4; ++j)
			w[j] = (uint32_t)t16[(seg * 8 + 2 * j) * S + c] | (uint32_t)t16[(seg * 8 + 2 * j + 1) * S + c] << 16;
		*reinterpret_cast<uint4*>(Xc + (long long)c * V + v0 + seg * 8) = make_uint4(w[0], w[1], w[2], w[3]);
	}
}

// Weight gradient on MFMA: dW[t][co][ci] = sum_v dZ[v][co] X[v + o_t][ci] as GEMMs with
// the voxels as the reduction, from channel-major copies (a lane's 8 reduction values are 8
// consecutive voxels along z of one channel, one 16-byte load). Voxel group g holds voxels
// 32 g .. 32 g + 31 (lane group kg: 8 g + ... one run of 8 along z, d2 % 8 = 0). A unit of
// work is (dx, dy, ci tile): the three 8-voxel pieces around the run in row (x + dx, y + dy)
// give the dz = -1, 0, +1 operands (the shifted ones by funnel shifts; pieces wrap along z).
// Wave w takes units w, w + 4, ...; each workgroup sums a contiguous range of groups in order
// and writes its partial sums in the layout k_reduce_chunks reduces: [t][co][ci], then the
// bias sums [co] (wave 0, lane groups reduced in kg order).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad_mfma(const uint16_t* __restrict__ dZc, const uint16_t* __restrict__ Xc, int d0, int d1,
                                                    int d2, long long groups_per_block, float* __restrict__ partial) {
	constexpr int TCI = CIN / 16, TCO = COUT / 16;
	constexpr int NU = 9 * TCI, UPW = (NU + 3) / 4;
	constexpr int NP = COUT * CIN;
	constexpr long long STRIDE = 27LL * NP + COUT;
	__shared__ float bsum[4][COUT];
	const long long V = (long long)d0 * d1 * d2;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int r16 = lane & 15, kg = lane >> 4;
	const int np8 = d2 / 8;
	f32x4 acc[UPW][3][TCO];
#pragma unroll
	for (int u = 0; u < UPW; ++u)
#pragma unroll
		for (int q = 0; q < 3; ++q)
#pragma unroll
			for (int n = 0; n < TCO; ++n) acc[u][q][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	float bacc[TCO];
#pragma unroll
	for (int n = 0; n < TCO; ++n) bacc[n] = 0.0f;
	const long long ng = V / 32;
	const long long g0 = (long long)blockIdx.x * groups_per_block;
	const long long g1 = g0 + groups_per_block < ng ? g0 + groups_per_block : ng;
	for (long long g = g0; g < g1; ++g) {
		const long long v0 = g * 32 + kg * 8;
		const int z = (int)(v0 % d2);
		const long long xy = v0 / d2;
		const int y = (int)(xy % d1), x = (int)(xy / d1);
		bf16x8 A[TCO];
#pragma unroll
		for (int n = 0; n < TCO; ++n) A[n] = load8(dZc + (long long)(n * 16 + r16) * V + v0);
		if (w == 0) {
#pragma unroll
			for (int n = 0; n < TCO; ++n) {
				const uint4 q = *reinterpret_cast<const uint4*>(dZc + (long long)(n * 16 + r16) * V + v0);
				float f[8];
				unpack8(q, f);
				float sum = 0.0f;
#pragma unroll
				for (int j = 0; j < 8; ++j) sum += f[j];
				bacc[n] += sum;
			}
		}
		const int p = z / 8, pm = p == 0 ? np8 - 1 : p - 1, pp = p == np8 - 1 ? 0 : p + 1;
#pragma unroll
		for (int u = 0; u < UPW; ++u) {
			const int unit = w + 4 * u;
			if (unit >= NU) break;
			const int tci = unit / 9, dxy = unit % 9;
			const int xx = wrap(x + dxy / 3 - 1, d0), yy = wrap(y + dxy % 3 - 1, d1);
			const uint16_t* xc = Xc + (long long)(tci * 16 + r16) * V + ((long long)xx * d1 + yy) * d2;
			const uint4 Lm = *reinterpret_cast<const uint4*>(xc + pm * 8);
			const uint4 L0 = *reinterpret_cast<const uint4*>(xc + p * 8);
			const uint4 Lp = *reinterpret_cast<const uint4*>(xc + pp * 8);
			bf16x8 B[3];
			B[0] = __builtin_bit_cast(bf16x8, make_uint4(__builtin_amdgcn_alignbit(L0.x, Lm.w, 16), __builtin_amdgcn_alignbit(L0.y, L0.x, 16),
			                                             __builtin_amdgcn_alignbit(L0.z, L0.y, 16), __builtin_amdgcn_alignbit(L0.w, L0.z, 16)));
			B[1] = __builtin_bit_cast(bf16x8, L0);
			B[2] = __builtin_bit_cast(bf16x8, make_uint4(__builtin_amdgcn_alignbit(L0.y, L0.x, 16), __builtin_amdgcn_alignbit(L0.z, L0.y, 16),
			                                             __builtin_amdgcn_alignbit(L0.w, L0.z, 16), __builtin_amdgcn_alignbit(Lp.x, L0.w, 16)));
#pragma unroll
			for (int q = 0; q < 3; ++q)
#pragma unroll
				for (int n = 0; n < TCO; ++n) acc[u][q][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[n], B[q], acc[u][q][n], 0, 0, 0);
		}
	}
	float* out = partial + (long long)blockIdx.x * STRIDE;
	// result layout: lane (kg, r16) holds D[co = n 16 + 4 kg + i][ci = tci 16 + r16]
#pragma unroll
	for (int u = 0; u < UPW; ++u) {
		const int unit = w + 4 * u;
		if (unit >= NU) break;
		const int tci = unit / 9, dxy = unit % 9;
#pragma unroll
		for (int q = 0; q < 3; ++q) {
			const int t = dxy * 3 + q;  // (dx + 1) 9 + (dy + 1) 3 + (dz + 1)
#pragma unroll
			for (int n = 0; n < TCO; ++n)
#pragma unroll
				for (int i = 0; i < 4; ++i)
					out[(long long)t * NP + (long long)(n * 16 + 4 * kg + i) * CIN + tci * 16 + r16] = acc[u][q][n][i];
		}
	}
	if (w == 0) {
#pragma unroll
		for (int n = 0; n < TCO; ++n) bsum[kg][n * 16 + r16] = bacc[n];
	}
	__syncthreads();
	if (threadIdx.x < COUT) out[27LL * NP + threadIdx.x] = bsum[0][threadIdx.x] + bsum[1][threadIdx.x] + bsum[2][threadIdx.x] + bsum[3][threadIdx.x];
}

// Weight gradient of the 16 -> 16 channel layers (the finest level, most of the training's
// FLOPs) on bricks staged in LDS: the operands of k_wgrad_mfma, but each workgroup loads its
// brick's rows once, 2 x 4 rows of 32 voxels along z, with the input's halo rows (4 x 6) and one
// 8-voxel piece of margin at each end, from the channel-major copies; the 27 taps' reads then
// come from LDS. Wave w takes brick rows 2w and 2w + 1 (one group of 32 voxels each) and all
// 9 (dx, dy) units; the waves' sums are added in wave order through LDS, and each workgroup
// writes one partial vector (taps, then the bias) for the fixed-order reduction. Needs
// d0 % 2 = d1 % 4 = d2 % 32 = 0.
constexpr int kWBX = 2, kWBY = 4;
__global__ void __launch_bounds__(256) k_wgrad_brick16(const uint16_t* __restrict__ dZc, const uint16_t* __restrict__ Xc, int d0, int d1,
                                                       int d2, int bricks_per_block, float* __restrict__ partial) {
	constexpr int C = 16, NP = C * C;
	constexpr int HX = kWBX + 2, HY = kWBY + 2, NR = kWBX * kWBY, NPC = 6;
	constexpr long long STRIDE = 27LL * NP + C;
	__shared__ uint4 xs[C * HX * HY * NPC];  // [ci][halo row][piece]: pieces z0 - 8 .. z0 + 39
	__shared__ uint4 zs[C * NR * 4];         // [co][row][piece]: z0 .. z0 + 31
	__shared__ float bsum[4][4][C];          // [wave][kg][co]
	const long long V = (long long)d0 * d1 * d2;
	const int np8 = d2 / 8, nbz = d2 / 32, nby = d1 / kWBY, nbx = d0 / kWBX;
	const int nbricks = nbx * nby * nbz;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int r16 = lane & 15, kg = lane >> 4;
	f32x4 acc[9][3];
#pragma unroll
	for (int u = 0; u < 9; ++u)
#pragma unroll
		for (int q = 0; q < 3; ++q) acc[u][q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	float bacc = 0.0f;
	const int b0 = blockIdx.x * bricks_per_block;
	const int b1 = b0 + bricks_per_block < nbricks ? b0 + bricks_per_block : nbricks;
	for (int b = b0; b < b1; ++b) {
		const int z0 = (b % nbz) * 32, y0 = ((b / nbz) % nby) * kWBY, x0 = (b / nbz / nby) * kWBX;
		__syncthreads();  // the previous brick's reads are done
		// one (channel, halo row) per thread and pass, its NPC pieces loaded together (only the
		// first and the last can wrap along z)
		const int zp = z0 / 8;
		for (int e = threadIdx.x; e < C * HX * HY; e += 256) {
			const int hrow = e % (HX * HY), ci = e / (HX * HY);
			const int gx = wrap(x0 - 1 + hrow / HY, d0), gy = wrap(y0 - 1 + hrow % HY, d1);
			const uint16_t* src = Xc + (long long)ci * V + ((long long)gx * d1 + gy) * d2;
			uint4 v[NPC];
#pragma unroll
			for (int pc = 0; pc < NPC; ++pc) {
				const int piece = pc == 0 ? (zp == 0 ? np8 - 1 : zp - 1) : (pc == NPC - 1 ? (zp + 4 == np8 ? 0 : zp + 4) : zp + pc - 1);
				v[pc] = *reinterpret_cast<const uint4*>(src + piece * 8);
			}
#pragma unroll
			for (int pc = 0; pc < NPC; ++pc) xs[e * NPC + pc] = v[pc];
		}
		if (threadIdx.x < C * NR) {
			const int row = threadIdx.x % NR, co = threadIdx.x / NR;
			const uint4* src = reinterpret_cast<const uint4*>(dZc + (long long)co * V + ((long long)(x0 + row / kWBY) * d1 + y0 + row % kWBY) * d2 + z0);
			uint4 v[4];
#pragma unroll
			for (int pc = 0; pc < 4; ++pc) v[pc] = src[pc];
#pragma unroll
			for (int pc = 0; pc < 4; ++pc) zs[threadIdx.x * 4 + pc] = v[pc];
		}
		__syncthreads();
#pragma unroll
		for (int rr = 0; rr < 2; ++rr) {
			const int row = 2 * w + rr, rx = row / kWBY, ry = row % kWBY;
			const uint4 aq = zs[(r16 * NR + row) * 4 + kg];
			const bf16x8 A = __builtin_bit_cast(bf16x8, aq);
			{
				float f[8];
				unpack8(aq, f);
#pragma unroll
				for (int j = 0; j < 8; ++j) bacc += f[j];
			}
#pragma unroll
			for (int u = 0; u < 9; ++u) {
				const int hrow = (rx + u / 3) * HY + ry + u % 3;  // (rx + 1 + dx, ry + 1 + dy)
				const int base = (r16 * HX * HY + hrow) * NPC + kg + 1;
				const uint4 Lm = xs[base - 1], L0 = xs[base], Lp = xs[base + 1];
				const bf16x8 Bm = __builtin_bit_cast(bf16x8, make_uint4(__builtin_amdgcn_alignbit(L0.x, Lm.w, 16), __builtin_amdgcn_alignbit(L0.y, L0.x, 16),
				                                                       __builtin_amdgcn_alignbit(L0.z, L0.y, 16), __builtin_amdgcn_alignbit(L0.w, L0.z, 16)));
				const bf16x8 Bc = __builtin_bit_cast(bf16x8, L0);
				const bf16x8 Bp = __builtin_bit_cast(bf16x8, make_uint4(__builtin_amdgcn_alignbit(L0.y, L0.x, 16), __builtin_amdgcn_alignbit(L0.z, L0.y, 16),
				                                                       __builtin_amdgcn_alignbit(L0.w, L0.z, 16), __builtin_amdgcn_alignbit(Lp.x, L0.w, 16)));
				acc[u][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bm, acc[u][0], 0, 0, 0);
				acc[u][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bc, acc[u][1], 0, 0, 0);
				acc[u][2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bp, acc[u][2], 0, 0, 0);
			}
		}
	}
	// the waves' sums, added in wave order in LDS (the staging buffer), then written coalesced
	__syncthreads();
	float* red = reinterpret_cast<float*>(xs);  // 27 x 256 floats
	for (int wv = 0; wv < 4; ++wv) {
		if (w == wv) {
#pragma unroll
			for (int u = 0; u < 9; ++u)
#pragma unroll
				for (int q = 0; q < 3; ++q)
#pragma unroll
					for (int i = 0; i < 4; ++i) {
						const int idx = (u * 3 + q) * NP + (4 * kg + i) * C + r16;  // [t][co][ci]
						red[idx] = (wv == 0 ? 0.0f : red[idx]) + acc[u][q][i];
					}
			bsum[wv][kg][r16] = bacc;
		}
		__syncthreads();
	}
	float* out = partial + (long long)blockIdx.x * STRIDE;
	for (int e = threadIdx.x; e < 27 * NP; e += 256) out[e] = red[e];
	if (threadIdx.x < C) {
		float sb = 0.0f;
		for (int wv = 0; wv < 4; ++wv)
			for (int k = 0; k < 4; ++k) sb += bsum[wv][k][threadIdx.x];
		out[27 * NP + threadIdx.x] = sb;
	}
}

// out[p] += sum_c partial[c][p], c in order
__global__ void k_reduce_chunks(const float* __restrict__ partial, int nchunk, long long stride, float* __restrict__ out) {
	long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= stride) return;
	float s = 0.0f;
	for (int c = 0; c < nchunk; ++c) s += partial[(long long)c * stride + p];
	out[p] += s;
}

// first pass of a long reduction: slice blockIdx.y of `per` chunks, in order, into tmp
__global__ void k_reduce_slice(const float* __restrict__ partial, int nchunk, int per, long long stride, float* __restrict__ tmp) {
	long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= stride) return;
	const int c0 = blockIdx.y * per, c1 = c0 + per < nchunk ? c0 + per : nchunk;
	float s = 0.0f;
	for (int c = c0; c < c1; ++c) s += partial[(long long)c * stride + p];
	tmp[(long long)blockIdx.y * stride + p] = s;
}

constexpr int kReduceSlices = 32;

// out[p] += the sum over the nchunk partial vectors: slices of consecutive chunks summed in
// order, then the slices in order (fixed order: deterministic)
void reduce_partials(const float* partial, int nchunk, long long stride, float* tmp, float* out, hipStream_t s) {
	const unsigned gx = (unsigned)((stride + 255) / 256);
	if (nchunk <= 2 * kReduceSlices) {
		hipLaunchKernelGGL(k_reduce_chunks, dim3(gx), dim3(256), 0, s, partial, nchunk, stride, out);
		return;
	}
	const int per = (nchunk + kReduceSlices - 1) / kReduceSlices;
	const int ns = (nchunk + per - 1) / per;
	hipLaunchKernelGGL(k_reduce_slice, dim3(gx, ns), dim3(256), 0, s, partial, nchunk, per, stride, tmp);
	hipLaunchKernelGGL(k_reduce_chunks, dim3(gx), dim3(256), 0, s, tmp, ns, stride, out);
}

int wgrad_chunk(long long V) {
	long long c = V / 32;
	c = (c + 63) / 64 * 64;
	return (int)(c < 64 ? 64 : c);
}

// Workgroups of the MFMA weight gradient: contiguous ranges of at least 16 groups of 32
// voxels, at most 1024 workgroups (256 for the widest layers), so that the partial sums stay
// small next to the work.
long long wgrad_blocks(long long V, long long stride, long long* per) {
	const long long ng = V / 32;
	const long long cap = stride <= 16384 ? 1024 : 256;
	const long long want = ng / 16 > 1 ? ng / 16 : 1;
	const long long target = want < cap ? want : cap;
	*per = (ng + target - 1) / target;
	return (ng + *per - 1) / *per;
}

bool wgrad_on_mfma(const int d[3]) {
	const long long V = (long long)d[0] * d[1] * d[2];
	return d[2] % 8 == 0 && V % 64 == 0;
}

// The weight gradient's kernel for a layer: the LDS brick kernel (16 -> 16 channels), the
// MFMA kernel, or the FP32 kernel; and the number of partial vectors it writes (one per
// workgroup, or per chunk), with `per` bricks, voxel groups or voxels per workgroup. The
// partial buffer is sized from the same count.
enum class WgradKind { Brick16, Mfma, Fp32 };

WgradKind wgrad_kind(int cin, int cout, const int d[3]) {
	if (!wgrad_on_mfma(d)) return WgradKind::Fp32;
	if (cin == 16 && cout == 16 && d[0] % kWBX == 0 && d[1] % kWBY == 0 && d[2] % 32 == 0) return WgradKind::Brick16;
	return WgradKind::Mfma;
}

long long wgrad_partials(int cin, int cout, const int d[3], long long* per) {
	const long long V = (long long)d[0] * d[1] * d[2];
	const long long stride = 27LL * cin * cout + cout;
	switch (wgrad_kind(cin, cout, d)) {
	case WgradKind::Brick16: {
		const long long nbricks = (long long)(d[0] / kWBX) * (d[1] / kWBY) * (d[2] / 32);
		*per = (nbricks + 1023) / 1024;
		return (nbricks + *per - 1) / *per;
	}
	case WgradKind::Mfma:
		return wgrad_blocks(V, stride, per);
	default:
		*per = wgrad_chunk(V);
		return (V + *per - 1) / *per;
	}
}

template <int CIN, int COUT>
void wgrad(const uint16_t* dZ, const uint16_t* X, const int d[3], float* partial, long long partial_len, float* grad_w,
           uint16_t* dzc, uint16_t* xc, float* tmp, hipStream_t s) {
	const long long V = (long long)d[0] * d[1] * d[2];
	const long long stride = 27LL * COUT * CIN + COUT;
	long long per = 0;
	const long long nblk = wgrad_partials(CIN, COUT, d, &per);
	if (nblk * stride > partial_len)
		throw std::runtime_error("learned preconditioner: weight-gradient partials exceed their buffer");
	const WgradKind kind = wgrad_kind(CIN, COUT, d);
	if (kind != WgradKind::Fp32) {
		hipLaunchKernelGGL((k_channel_major<COUT>), dim3((unsigned)(V / 64)), dim3(256), 0, s, dZ, V, dzc);
		hipLaunchKernelGGL((k_channel_major<CIN>), dim3((unsigned)(V / 64)), dim3(256), 0, s, X, V, xc);
	}
	if (kind == WgradKind::Brick16)
		hipLaunchKernelGGL(k_wgrad_brick16, dim3((unsigned)nblk), dim3(256), 0, s, dzc, xc, d[0], d[1], d[2], (int)per, partial);
	else if (kind == WgradKind::Mfma)
		hipLaunchKernelGGL((k_wgrad_mfma<CIN, COUT>), dim3((unsigned)nblk), dim3(256), 0, s, dzc, xc, d[0], d[1], d[2], per, partial);
	else
		hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(28, (unsigned)nblk), dim3(256), 0, s, dZ, X, d[0], d[1], d[2], (int)per, partial);
	reduce_partials(partial, (int)nblk, stride, tmp, grad_w, s);
}

void wgrad_any(int cin, int cout, const uint16_t* dZ, const uint16_t* X, const int d[3], float* partial, long long partial_len,
               float* grad_w, uint16_t* dzc, uint16_t* xc, float* tmp, hipStream_t s) {
	if (cin == 16 && cout == 16) return wgrad<16, 16>(dZ, X, d, partial, partial_len, grad_w, dzc, xc, tmp, s);
	if (cin == 16 && cout == 32) return wgrad<16, 32>(dZ, X, d, partial, partial_len, grad_w, dzc, xc, tmp, s);
	if (cin == 32 && cout == 64) return wgrad<32, 64>(dZ, X, d, partial, partial_len, grad_w, dzc, xc, tmp, s);
	if (cin == 64 && cout == 32) return wgrad<64, 32>(dZ, X, d, partial, partial_len, grad_w, dzc, xc, tmp, s);
	if (cin == 32 && cout == 16) return wgrad<32, 16>(dZ, X, d, partial, partial_len, grad_w, dzc, xc, tmp, s);
	throw std::runtime_error("wgrad: unsupported channel pair");
}

// ------------------------------------------------------------------ elementwise pieces
__global__ void k_sumsq(const float* __restrict__ a, long long n, double* __restrict__ partial) {
	__shared__ double red[256];
	double s = 0.0;
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
		s += (double)a[i] * a[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sumsq_diff(const float* __restrict__ a, const float* __restrict__ b, long long n, double* __restrict__ partial) {
	__shared__ double red[256];
	double s = 0.0;
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
		double d = (double)a[i] - b[i];
		s += d * d;
	}
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[k] = sum of partial[]; mode 0: plain, 1: s = sqrt(sum / n3) into out[1] too, 2: loss = out[3] / out[2] into out[4]
__global__ void k_reduce_scalar(const double* __restrict__ partial, int n, double* __restrict__ out, int k, int mode, double n3) {
	__shared__ double red[256];
	double s = 0.0;
	for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		out[k] = red[0];
		if (mode == 1) out[1] = sqrt(red[0] / n3);
		if (mode == 2) out[4] += out[2] > 0 ? red[0] / out[2] : 0.0;
	}
}

// x0[v][0..2] = r / s, [3..11] = mu / mu_scale, [12..15] = 0
__global__ void k_pack_input(const float* __restrict__ r, const float* __restrict__ mu, const double* __restrict__ sc,
                             float inv_mu_scale, long long V, uint16_t* __restrict__ x0) {
	long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= V) return;
	double s = sc[1];
	float is = s > 0 ? (float)(1.0 / s) : 0.0f;
	uint16_t out[16];
#pragma unroll
	for (int c = 0; c < 3; ++c) out[c] = f2bf(r[c * V + v] * is);
#pragma unroll
	for (int c = 0; c < 9; ++c) out[3 + c] = f2bf(mu[c * V + v] * inv_mu_scale);
#pragma unroll
	for (int c = 12; c < 16; ++c) out[c] = 0;
	uint4* dst = reinterpret_cast<uint4*>(x0 + v * 16);
	dst[0] = make_uint4(out[0] | (uint32_t)out[1] << 16, out[2] | (uint32_t)out[3] << 16, out[4] | (uint32_t)out[5] << 16,
	                    out[6] | (uint32_t)out[7] << 16);
	dst[1] = make_uint4(out[8] | (uint32_t)out[9] << 16, out[10] | (uint32_t)out[11] << 16, out[12] | (uint32_t)out[13] << 16,
	                    out[14] | (uint32_t)out[15] << 16);
}

// mean over the 8 children (bf16 in/out); one thread per coarse voxel and 8 channels, the
// children summed in the order (a, b, q) of the three axes
__global__ void k_pool(const uint16_t* __restrict__ x, int c, int d0, int d1, int d2, uint16_t* __restrict__ y) {
	const int h0 = d0 / 2, h1 = d1 / 2, h2 = d2 / 2, cv = c / 8;
	long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	long long n = (long long)h0 * h1 * h2 * cv;
	if (e >= n) return;
	const int q = (int)(e % cv);
	const int v = (int)(e / cv);
	const int z = v % h2, yy = (v / h2) % h1, xx = v / (h1 * h2);
	float s[8] = {};
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int r = 0; r < 2; ++r) {
				float f[8];
				unpack8(reinterpret_cast<const uint4*>(x + (((long long)(2 * xx + a) * d1 + 2 * yy + b) * d2 + 2 * z + r) * c)[q], f);
#pragma unroll
				for (int i = 0; i < 8; ++i) s[i] += f[i];
			}
#pragma unroll
	for (int i = 0; i < 8; ++i) s[i] *= 0.125f;
	reinterpret_cast<uint4*>(y + (long long)v * c)[q] = pack8(s);
}

// u = up(coarse) + skip (bf16); fine dims d; one thread per fine voxel and 8 channels
__global__ void k_up_add(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip, int c, int d0, int d1, int d2,
                         uint16_t* __restrict__ u) {
	const int cv = c / 8;
	long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	long long n = (long long)d0 * d1 * d2 * cv;
	if (e >= n) return;
	const int q = (int)(e % cv);
	const long long v = e / cv;
	const int z = (int)(v % d2), yy = (int)((v / d2) % d1), xx = (int)(v / ((long long)d1 * d2));
	const long long pv = ((long long)(xx / 2) * (d1 / 2) + yy / 2) * (d2 / 2) + z / 2;
	float a[8], b[8];
	unpack8(reinterpret_cast<const uint4*>(coarse + pv * c)[q], a);
	unpack8(reinterpret_cast<const uint4*>(skip + v * c)[q], b);
#pragma unroll
	for (int i = 0; i < 8; ++i) a[i] += b[i];
	reinterpret_cast<uint4*>(u + v * c)[q] = pack8(a);
}

// z[c][v] = r[c][v] + s (sum_ci d1[v][ci] Wh[c][ci] + bh[c])
__global__ void k_head(const uint16_t* __restrict__ d1, const float* __restrict__ Wh, const float* __restrict__ bh,
                       const float* __restrict__ r, const double* __restrict__ sc, long long V, float* __restrict__ z) {
	long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= V) return;
	float s = (float)sc[1];
	float x[kHeadIn];
#pragma unroll
	for (int i = 0; i < kHeadIn; ++i) x[i] = bf2f(d1[v * kHeadIn + i]);
#pragma unroll
	for (int c = 0; c < kHeadOut; ++c) {
		float y = bh[c];
#pragma unroll
		for (int i = 0; i < kHeadIn; ++i) y += x[i] * Wh[c * kHeadIn + i];
		z[c * V + v] = r[c * V + v] + s * y;
	}
}

// dy = s 2 (z - e) / ||e||^2; dd1[v][i] = sum_c dy[c] Wh[c][i]; head parameter partial sums
__global__ void __launch_bounds__(256) k_head_backward(const uint16_t* __restrict__ d1, const float* __restrict__ Wh,
                                                       const float* __restrict__ z, const float* __restrict__ e,
                                                       const double* __restrict__ sc, long long V, float* __restrict__ dd1,
                                                       float* __restrict__ partial) {
	constexpr int NQ = kHeadOut * kHeadIn + kHeadOut;
	__shared__ float red[NQ][64];
	float acc[NQ];
#pragma unroll
	for (int q = 0; q < NQ; ++q) acc[q] = 0.0f;
	const float f = sc[2] > 0 ? (float)(sc[1] * 2.0 / sc[2]) : 0.0f;
	for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long long)gridDim.x * blockDim.x) {
		float dy[kHeadOut];
#pragma unroll
		for (int c = 0; c < kHeadOut; ++c) dy[c] = f * (z[c * V + v] - e[c * V + v]);
#pragma unroll
		for (int i = 0; i < kHeadIn; ++i) {
			float x = bf2f(d1[v * kHeadIn + i]);
			float g = 0.0f;
#pragma unroll
			for (int c = 0; c < kHeadOut; ++c) {
				g += dy[c] * Wh[c * kHeadIn + i];
				acc[c * kHeadIn + i] += dy[c] * x;
			}
			dd1[v * kHeadIn + i] = g;
		}
#pragma unroll
		for (int c = 0; c < kHeadOut; ++c) acc[kHeadOut * kHeadIn + c] += dy[c];
	}
	// block reduction: lanes of each wave, then the 4 waves in order
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for (int q = 0; q < NQ; ++q) {
		float v = acc[q];
		for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
		if (lane == 0) red[q][wave] = v;
	}
	__syncthreads();
	if (threadIdx.x < NQ) partial[(long long)blockIdx.x * NQ + threadIdx.x] = red[threadIdx.x][0] + red[threadIdx.x][1] +
	                                                                          red[threadIdx.x][2] + red[threadIdx.x][3];
}

// dZ = bf16(dout * (act > 0)); 8 values per thread (n % 8 = 0)
__global__ void k_relu_mask(const float* __restrict__ dout, const uint16_t* __restrict__ act, long long n, uint16_t* __restrict__ dz) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n / 8) return;
	const float4 g0 = reinterpret_cast<const float4*>(dout)[2 * i], g1 = reinterpret_cast<const float4*>(dout)[2 * i + 1];
	float a[8];
	unpack8(reinterpret_cast<const uint4*>(act)[i], a);
	const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
	float o[8];
#pragma unroll
	for (int j = 0; j < 8; ++j) o[j] = a[j] > 0.0f ? g[j] : 0.0f;
	reinterpret_cast<uint4*>(dz)[i] = pack8(o);
}

// coarse[p][c] = sum of the 8 children of fine[.][c]
__global__ void k_pool_sum(const float* __restrict__ fine, int c, int d0, int d1, int d2, float* __restrict__ coarse) {
	const int h0 = d0 / 2, h1 = d1 / 2, h2 = d2 / 2;
	long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	long long n = (long long)h0 * h1 * h2 * c;
	if (e >= n) return;
	int ch = (int)(e % c);
	long long v = e / c;
	int z = (int)(v % h2), yy = (int)((v / h2) % h1), xx = (int)(v / ((long long)h1 * h2));
	float s = 0.0f;
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int q = 0; q < 2; ++q) s += fine[(((long long)(2 * xx + a) * d1 + 2 * yy + b) * d2 + 2 * z + q) * c + ch];
	coarse[e] = s;
}

// out = up(coarse) / 8 + skip (float32); fine dims d; one thread per fine voxel and 4 channels
__global__ void k_unpool_add(const float* __restrict__ coarse, const float* __restrict__ skip, int c, int d0, int d1, int d2,
                             float* __restrict__ out) {
	const int c4 = c / 4;
	long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	long long n = (long long)d0 * d1 * d2 * c4;
	if (e >= n) return;
	const int q = (int)(e % c4);
	const long long v = e / c4;
	const int z = (int)(v % d2), yy = (int)((v / d2) % d1), xx = (int)(v / ((long long)d1 * d2));
	const long long pv = ((long long)(xx / 2) * (d1 / 2) + yy / 2) * (d2 / 2) + z / 2;
	const float4 a = reinterpret_cast<const float4*>(coarse + pv * c)[q], b = reinterpret_cast<const float4*>(skip + v * c)[q];
	reinterpret_cast<float4*>(out + v * c)[q] =
	    make_float4(a.x * 0.125f + b.x, a.y * 0.125f + b.y, a.z * 0.125f + b.z, a.w * 0.125f + b.w);
}

// grad[q] += sum over the nblk block partials of value q: one workgroup per q, a fixed
// lane order and a fixed tree (deterministic)
__global__ void __launch_bounds__(256) k_reduce_head(const float* __restrict__ partial, int nblk, float* __restrict__ grad) {
	constexpr int NQ = kHeadOut * kHeadIn + kHeadOut;
	__shared__ float red[256];
	const int q = blockIdx.x;
	float s = 0.0f;
	for (int b = threadIdx.x; b < nblk; b += 256) s += partial[b * NQ + q];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) grad[q] += red[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, long long n,
                       float lr, float inv_batch, float c1, float c2) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	float gi = g[i] * inv_batch;
	float mi = b1 * m[i] + (1.0f - b1) * gi;
	float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 forward weights W[t][co][ci] and adjoint weights W'[t][ci][co] = W[26 - t][co][ci]
__global__ void k_weights_bf16(const float* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf, uint16_t* __restrict__ wb) {
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	int n = 27 * cin * cout;
	if (e >= n) return;
	int ci = e % cin, co = (e / cin) % cout, t = e / (cin * cout);
	uint16_t h = f2bf(w[e]);
	wf[e] = h;
	wb[((26 - t) * cin + ci) * cout + co] = h;
}

inline unsigned nb(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <class T>
T* dmalloc(long long n) {
	void* p = nullptr;
	PC_CHECK(hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(T)));
	return static_cast<T*>(p);
}

}  // namespace

// ------------------------------------------------------------------ UNet
UNet::UNet(const int d[3]) {
	for (int a = 0; a < 3; ++a) {
		if (d[a] < 4 || d[a] % 4) throw std::runtime_error("UNet: every grid dimension must be a multiple of 4");
		d_[a] = d[a];
	}
	V_ = (long long)d[0] * d[1] * d[2];
	const int spec[6][3] = {{16, 16, 0}, {16, 32, 1}, {32, 64, 2}, {64, 32, 2}, {32, 16, 1}, {16, 16, 0}};
	long long off = 0;
	for (auto& sp : spec) {
		ConvSpec c{sp[0], sp[1], sp[2], off, off + 27LL * sp[0] * sp[1]};
		off = c.b_off + sp[1];
		convs_.push_back(c);
	}
	long long n = param_count();
	params_ = dmalloc<float>(n);
	grads_ = dmalloc<float>(n);
	m_ = dmalloc<float>(n);
	v_ = dmalloc<float>(n);
	PC_CHECK(hipMemset(grads_, 0, n * sizeof(float)));
	PC_CHECK(hipMemset(m_, 0, n * sizeof(float)));
	PC_CHECK(hipMemset(v_, 0, n * sizeof(float)));
	long long nw = 0;
	for (auto& c : convs_) nw += 27LL * c.cin * c.cout;
	wf_ = dmalloc<uint16_t>(nw);
	wb_ = dmalloc<uint16_t>(nw);
	const long long V = V_, V1 = V / 8, V2 = V / 64;
	x0_ = dmalloc<uint16_t>(16 * V);
	e1_ = dmalloc<uint16_t>(16 * V);
	p1_ = dmalloc<uint16_t>(16 * V1);
	e2_ = dmalloc<uint16_t>(32 * V1);
	p2_ = dmalloc<uint16_t>(32 * V2);
	e3_ = dmalloc<uint16_t>(64 * V2);
	d3_ = dmalloc<uint16_t>(32 * V2);
	u2_ = dmalloc<uint16_t>(32 * V1);
	d2_ = dmalloc<uint16_t>(16 * V1);
	u1_ = dmalloc<uint16_t>(16 * V);
	d1_ = dmalloc<uint16_t>(16 * V);
	dz_ = dmalloc<uint16_t>(16 * V);
	g0_ = dmalloc<float>(16 * V);
	g0b_ = dmalloc<float>(16 * V);
	g1_ = dmalloc<float>(32 * V1);
	g1b_ = dmalloc<float>(32 * V1);
	g2_ = dmalloc<float>(64 * V2);
	z_ = dmalloc<float>(3 * V);
	long long worst = 0;
	for (auto& c : convs_) {
		const long long stride = 27LL * c.cin * c.cout + c.cout;
		const int dl[3] = {d[0] >> c.level, d[1] >> c.level, d[2] >> c.level};
		long long per = 0;
		const long long need = wgrad_partials(c.cin, c.cout, dl, &per) * stride;
		if (need > worst) worst = need;
	}
	dzc_ = dmalloc<uint16_t>(16 * V);  // channel-major copies for the weight gradients (largest: level 0, 16 channels)
	xc_ = dmalloc<uint16_t>(16 * V);
	rtmp_ = dmalloc<float>((long long)kReduceSlices * (27LL * 64 * 32 + 64));
	long long head = (long long)kRedBlocks * (kHeadOut * kHeadIn + kHeadOut);
	partial_len_ = worst > head ? worst : head;
	partial_ = dmalloc<float>(partial_len_);
	dscal_ = dmalloc<double>(8);
	dpart_ = dmalloc<double>(kRedBlocks);
	PC_CHECK(hipMemset(dscal_, 0, 8 * sizeof(double)));
	// He-normal convolution weights, zero biases and a zero head: the untrained network is z = r
	std::vector<float> h(n, 0.0f);
	std::mt19937_64 rng(12345);
	for (auto& c : convs_) {
		std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * c.cin)));
		for (long long i = 0; i < 27LL * c.cin * c.cout; ++i) h[c.w_off + i] = nd(rng);
	}
	set_params(h.data(), nullptr);
	PC_CHECK(hipDeviceSynchronize());
}

UNet::~UNet() {
	for (void* p : {(void*)params_, (void*)grads_, (void*)m_, (void*)v_, (void*)wf_, (void*)wb_, (void*)x0_, (void*)e1_, (void*)p1_,
	                (void*)e2_, (void*)p2_, (void*)e3_, (void*)d3_, (void*)u2_, (void*)d2_, (void*)u1_, (void*)d1_, (void*)dz_,
	                (void*)g0_, (void*)g0b_, (void*)g1_, (void*)g1b_, (void*)g2_, (void*)z_, (void*)partial_, (void*)dscal_,
	                (void*)dpart_, (void*)dzc_, (void*)xc_, (void*)rtmp_})
		if (p) (void)hipFree(p);
}

long long UNet::param_count() {
	const int spec[6][2] = {{16, 16}, {16, 32}, {32, 64}, {64, 32}, {32, 16}, {16, 16}};
	long long n = 0;
	for (auto& sp : spec) n += 27LL * sp[0] * sp[1] + sp[1];
	return n + kHeadOut * kHeadIn + kHeadOut;
}

void UNet::refresh_weights(hipStream_t s) {
	long long off = 0;
	for (auto& c : convs_) {
		int n = 27 * c.cin * c.cout;
		hipLaunchKernelGGL(k_weights_bf16, dim3(nb(n, 256)), dim3(256), 0, s, params_ + c.w_off, c.cin, c.cout, wf_ + off, wb_ + off);
		off += n;
	}
}

void UNet::set_params(const float* host, hipStream_t s) {
	PC_CHECK(hipMemcpyAsync(params_, host, param_count() * sizeof(float), hipMemcpyHostToDevice, s));
	refresh_weights(s);
	PC_CHECK(hipStreamSynchronize(s));
}

void UNet::get_params(float* host, hipStream_t s) const {
	PC_CHECK(hipMemcpyAsync(host, params_, param_count() * sizeof(float), hipMemcpyDeviceToHost, s));
	PC_CHECK(hipStreamSynchronize(s));
}

void UNet::get_grads(float* host, hipStream_t s) const {
	PC_CHECK(hipMemcpyAsync(host, grads_, param_count() * sizeof(float), hipMemcpyDeviceToHost, s));
	PC_CHECK(hipStreamSynchronize(s));
}

void UNet::zero_grad(hipStream_t s) {
	PC_CHECK(hipMemsetAsync(grads_, 0, param_count() * sizeof(float), s));
	PC_CHECK(hipMemsetAsync(dscal_ + 4, 0, sizeof(double), s));
}

static long long weight_offset(const std::vector<ConvSpec>& convs, size_t k) {
	long long off = 0;
	for (size_t i = 0; i < k; ++i) off += 27LL * convs[i].cin * convs[i].cout;
	return off;
}

void UNet::conv_forward(const ConvSpec& c, const uint16_t* x, uint16_t* y, int lvl, hipStream_t s) {
	const int dl[3] = {d_[0] >> lvl, d_[1] >> lvl, d_[2] >> lvl};
	size_t k = &c - convs_.data();
	conv3_any<uint16_t>(c.cin, c.cout, x, wf_ + weight_offset(convs_, k), params_ + c.b_off, true, y, dl, s);
}

// dout: gradient of the layer's output (float32, [V][cout]); act: its output; input: its
// input; din (optional): gradient of the input (float32 [V][cin]).
void UNet::conv_backward(const ConvSpec& c, const float* dout, const uint16_t* act, const uint16_t* input, float* din, int lvl,
                         hipStream_t s) {
	const int dl[3] = {d_[0] >> lvl, d_[1] >> lvl, d_[2] >> lvl};
	long long Vl = level_voxels(lvl);
	hipLaunchKernelGGL(k_relu_mask, dim3(nb(Vl * c.cout / 8, 256)), dim3(256), 0, s, dout, act, Vl * c.cout, dz_);
	wgrad_any(c.cin, c.cout, dz_, input, dl, partial_, partial_len_, grads_ + c.w_off, dzc_, xc_, rtmp_, s);
	if (din) {
		size_t k = &c - convs_.data();
		conv3_any<float>(c.cout, c.cin, dz_, wb_ + weight_offset(convs_, k), nullptr, false, din, dl, s);
	}
}

void UNet::forward(const float* r, const float* mu, float mu_scale, float* z, hipStream_t s) {
	const long long V = V_;
	hipLaunchKernelGGL(k_sumsq, dim3(kRedBlocks), dim3(256), 0, s, r, 3 * V, dpart_);
	hipLaunchKernelGGL(k_reduce_scalar, dim3(1), dim3(256), 0, s, dpart_, kRedBlocks, dscal_, 0, 1, (double)(3 * V));
	hipLaunchKernelGGL(k_pack_input, dim3(nb(V, 256)), dim3(256), 0, s, r, mu, dscal_, mu_scale > 0 ? 1.0f / mu_scale : 0.0f, V, x0_);
	const int* d = d_;
	conv_forward(convs_[0], x0_, e1_, 0, s);
	hipLaunchKernelGGL(k_pool, dim3(nb(2 * V / 8, 256)), dim3(256), 0, s, e1_, 16, d[0], d[1], d[2], p1_);
	conv_forward(convs_[1], p1_, e2_, 1, s);
	hipLaunchKernelGGL(k_pool, dim3(nb(4 * V / 64, 256)), dim3(256), 0, s, e2_, 32, d[0] / 2, d[1] / 2, d[2] / 2, p2_);
	conv_forward(convs_[2], p2_, e3_, 2, s);
	conv_forward(convs_[3], e3_, d3_, 2, s);
	hipLaunchKernelGGL(k_up_add, dim3(nb(4 * V / 8, 256)), dim3(256), 0, s, d3_, e2_, 32, d[0] / 2, d[1] / 2, d[2] / 2, u2_);
	conv_forward(convs_[4], u2_, d2_, 1, s);
	hipLaunchKernelGGL(k_up_add, dim3(nb(2 * V, 256)), dim3(256), 0, s, d2_, e1_, 16, d[0], d[1], d[2], u1_);
	conv_forward(convs_[5], u1_, d1_, 0, s);
	const long long hw = convs_.back().b_off + convs_.back().cout;
	hipLaunchKernelGGL(k_head, dim3(nb(V, 256)), dim3(256), 0, s, d1_, params_ + hw, params_ + hw + kHeadOut * kHeadIn, r, dscal_, V,
	                   z);
}

void UNet::accumulate_gradient(const float* r, const float* mu, float mu_scale, const float* e, hipStream_t s) {
	const long long V = V_;
	const int* d = d_;
	forward(r, mu, mu_scale, z_, s);
	hipLaunchKernelGGL(k_sumsq, dim3(kRedBlocks), dim3(256), 0, s, e, 3 * V, dpart_);
	hipLaunchKernelGGL(k_reduce_scalar, dim3(1), dim3(256), 0, s, dpart_, kRedBlocks, dscal_, 2, 0, 0.0);
	hipLaunchKernelGGL(k_sumsq_diff, dim3(kRedBlocks), dim3(256), 0, s, z_, e, 3 * V, dpart_);
	hipLaunchKernelGGL(k_reduce_scalar, dim3(1), dim3(256), 0, s, dpart_, kRedBlocks, dscal_, 3, 2, 0.0);
	const long long hw = convs_.back().b_off + convs_.back().cout;
	hipLaunchKernelGGL(k_head_backward, dim3(kRedBlocks), dim3(256), 0, s, d1_, params_ + hw, z_, e, dscal_, V, g0_, partial_);
	hipLaunchKernelGGL(k_reduce_head, dim3(kHeadOut * kHeadIn + kHeadOut), dim3(256), 0, s, partial_, kRedBlocks, grads_ + hw);
	const ConvSpec &enc1 = convs_[0], &enc2 = convs_[1], &enc3 = convs_[2], &dec3 = convs_[3], &dec2 = convs_[4], &dec1 = convs_[5];
	conv_backward(dec1, g0_, d1_, u1_, g0b_, 0, s);  // du1 -> g0b (also the skip gradient of e1)
	hipLaunchKernelGGL(k_pool_sum, dim3(nb(16 * V / 8, 256)), dim3(256), 0, s, g0b_, 16, d[0], d[1], d[2], g1_);
	conv_backward(dec2, g1_, d2_, u2_, g1b_, 1, s);  // du2 -> g1b (skip gradient of e2)
	hipLaunchKernelGGL(k_pool_sum, dim3(nb(32 * V / 64, 256)), dim3(256), 0, s, g1b_, 32, d[0] / 2, d[1] / 2, d[2] / 2, g2_);
	conv_backward(dec3, g2_, d3_, e3_, g2_, 2, s);   // de3 -> g2 (dz was taken from g2 first)
	conv_backward(enc3, g2_, e3_, p2_, g2_, 2, s);   // dp2 -> g2
	hipLaunchKernelGGL(k_unpool_add, dim3(nb(8 * V / 8, 256)), dim3(256), 0, s, g2_, g1b_, 32, d[0] / 2, d[1] / 2, d[2] / 2, g1_);
	conv_backward(enc2, g1_, e2_, p1_, g1_, 1, s);   // dp1 -> g1
	hipLaunchKernelGGL(k_unpool_add, dim3(nb(4 * V, 256)), dim3(256), 0, s, g1_, g0b_, 16, d[0], d[1], d[2], g0_);
	conv_backward(enc1, g0_, e1_, x0_, nullptr, 0, s);
}

void UNet::adam_step(float lr, int batch, hipStream_t s) {
	++adam_t_;
	float c1 = 1.0f - std::pow(0.9f, (float)adam_t_), c2 = 1.0f - std::pow(0.999f, (float)adam_t_);
	long long n = param_count();
	hipLaunchKernelGGL(k_adam, dim3(nb(n, 256)), dim3(256), 0, s, params_, grads_, m_, v_, n, lr, 1.0f / (float)(batch > 0 ? batch : 1),
	                   c1, c2);
	refresh_weights(s);
}

}  // namespace picgpu
