// FP32 particle kernels (the chosen form): positions kept as an integer cell plus an FP32
// offset in [0, 1) per axis, momentum as u = p / (m c) = gamma v in FP32.
//
// The algorithm is the reference's (same predictor, Picard iterations, alpha tensor and
// CIC maps, PIC++/mover.cpp:194-491), restated on cell offsets: the node-map weights of a
// particle at offset f are (1 - f, f) from its cell's lower node, and the cell-map weights
// are (1/2 - f, 1/2 + f) from the cell below when f < 1/2, else (3/2 - f, f - 1/2) from its
// own cell — the closed forms of correlationBspline (PIC++/correlation.cpp:824-835).
//
// Deposits are deterministic without atomics. Particles are sorted by key (species, cell,
// octant); a group of G lanes serves one cell (node moments) or one dual cell (cell moments,
// density) and takes its particles round-robin. Each lane sums its particles in that fixed
// order, the group then reduces through LDS in a fixed lane order and writes 8 corner sums.
#include <type_traits>

#include "pic_device.hpp"
#include "pic_kernels.hpp"
#include "pic_kernels_fp32.hpp"
#include "pic_tiles.hpp"

namespace picgpu {

constexpr float kPiF = 3.14159265358979323846f;

struct Map32 {
	int nb, cb;    // node-map and cell-map base index (ghosted)
	float nw[2];   // node weights
	float cw[2];   // cell weights
};

__device__ inline Map32 map32(int c, float f) {
	Map32 m;
	m.nb = c;
	m.nw[0] = 1.0f - f;
	m.nw[1] = f;
	if (f < 0.5f) {
		m.cb = c - 1;
		m.cw[0] = 0.5f - f;
		m.cw[1] = 0.5f + f;
	} else {
		m.cb = c;
		m.cw[0] = 1.5f - f;
		m.cw[1] = f - 0.5f;
	}
	return m;
}

__device__ inline void cell_ijk(const Grid& g, int cell, int c[3]) {
	c[2] = cell % g.N[2];
	c[1] = (cell / g.N[2]) % g.N[1];
	c[0] = cell / (g.N[2] * g.N[1]);
}

// Maps at a point given relative to cell c (offsets may leave [0, 1) during the Picard
// iterations); bases are clamped into the ghosted arrays, returning false if a clamp was needed.
__device__ inline bool maps32(const Grid& g, const int c[3], const float f[3], Map32 m[3]) {
	bool ok = true;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		float fl = floorf(f[a]);
		int cc = c[a] + (int)fl;
		m[a] = map32(cc, f[a] - fl);
		if (m[a].nb < 0 || m[a].nb > g.N[a] - 1) {
			m[a].nb = m[a].nb < 0 ? 0 : g.N[a] - 1;
			ok = false;
		}
		if (m[a].cb < 0 || m[a].cb > g.N[a] - 2) {
			m[a].cb = m[a].cb < 0 ? 0 : g.N[a] - 2;
			ok = false;
		}
	}
	return ok;
}

// FP32 field copies are node-major float4 (x, y, z, 0): one 16-byte load per stencil point.
__device__ inline void gather_node32(const Grid& g, const float4* __restrict__ f, const Map32* m, float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
	const int sj = g.N[2] + 1, si = (g.N[1] + 1) * sj;
	const int base = m[0].nb * si + m[1].nb * sj + m[2].nb;
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j)
#pragma unroll
			for (int k = 0; k < 2; ++k) {
				float4 e = f[base + i * si + j * sj + k];
				float c = m[0].nw[i] * m[1].nw[j] * m[2].nw[k];
				out[0] += e.x * c;
				out[1] += e.y * c;
				out[2] += e.z * c;
			}
}

__device__ inline void gather_cell32(const Grid& g, const float4* __restrict__ f, const Map32* m, float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
	const int sj = g.N[2], si = g.N[1] * sj;
	const int base = m[0].cb * si + m[1].cb * sj + m[2].cb;
#pragma unroll
	for (int i = 0; i < 2; ++i)
#pragma unroll
		for (int j = 0; j < 2; ++j)
#pragma unroll
			for (int k = 0; k < 2; ++k) {
				float4 b = f[base + i * si + j * sj + k];
				float c = m[0].cw[i] * m[1].cw[j] * m[2].cw[k];
				out[0] += b.x * c;
				out[1] += b.y * c;
				out[2] += b.z * c;
			}
}

// The push's field boxes in LDS. A workgroup's 256 particles are consecutive in the sorted
// order; when they all lie in one (i, j) row of cells, k0 <= k <= k0 + K - 1 with K <= kBoxK,
// the workgroup stages the fields around that row: E^n on the row's own nodes (the start
// position's node map), E^{n+theta} with one node and B with two cells more on every side
// (the maps of Picard iterates up to one cell away). A gather whose 2 x 2 x 2 stencil lies
// in its box reads LDS, any other reads HBM; both sum the same values in the same order.
// Vector memory instructions (one per stencil point and field) bound the push otherwise.
constexpr int kBoxK = 20;

// LDS pointers carry their address space, so that the box and HBM paths of a gather stay
// ds_read and global_load instructions (a generic pointer would make both flat loads).
typedef const __attribute__((address_space(3))) float lds_float;

// A box W x W x (K + W - 1) points, laid out with the strides of the largest K (kBoxK) so
// that a stencil's 8 LDS addresses differ by constants.
template <int W>
struct Box32 {
	static constexpr int sj = kBoxK + W - 1, si = W * sj, size = W * si;
	int o[3];  // ghosted index of local (0, 0, 0)
	int limk;  // largest local stencil base along k: K + W - 3
};

template <int W>
__device__ inline Box32<W> make_box(int i0, int j0, int k0, int K) {
	Box32<W> b;
	constexpr int margin = W == 2 ? 0 : W - 3;
	b.o[0] = i0 - margin;
	b.o[1] = j0 - margin;
	b.o[2] = k0 - margin;
	b.limk = K + W - 3;
	return b;
}

// local base of the stencil at ghosted base (a, c, d); false when the stencil leaves the box
template <int W>
__device__ inline bool in_box(const Box32<W>& b, int a, int c, int d, int& l) {
	a -= b.o[0];
	c -= b.o[1];
	d -= b.o[2];
	l = a * Box32<W>::si + c * Box32<W>::sj + d;
	return (unsigned)a <= (unsigned)(W - 2) && (unsigned)c <= (unsigned)(W - 2) && (unsigned)d <= (unsigned)b.limk;
}

__device__ inline void sum8(const float e[8][3], const float w0[2], const float w1[2], const float w2[2], float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		float c = w0[(t >> 2) & 1] * w1[(t >> 1) & 1] * w2[t & 1];
		out[0] += e[t][0] * c;
		out[1] += e[t][1] * c;
		out[2] += e[t][2] * c;
	}
}

// 8-point gather of a node (NODE) or cell field with the maps m, from the LDS box when the
// workgroup has one (use) and the stencil lies in it, else from HBM. Same sums as
// gather_node32 / gather_cell32.
template <bool NODE, int W>
__device__ inline void gather_boxed(const Grid& g, const float4* __restrict__ f, lds_float* box, const Box32<W>& b, bool use,
                                   const Map32* m, float out[3]) {
	const int a = NODE ? m[0].nb : m[0].cb, c = NODE ? m[1].nb : m[1].cb, d = NODE ? m[2].nb : m[2].cb;
	const float* w0 = NODE ? m[0].nw : m[0].cw;
	const float* w1 = NODE ? m[1].nw : m[1].cw;
	const float* w2 = NODE ? m[2].nw : m[2].cw;
	int l;
	float e[8][3];
	if (use && in_box(b, a, c, d, l)) {
		lds_float* p0 = box + 4 * l;  // the 8 points at constant offsets from it
#pragma unroll
		for (int t = 0; t < 8; ++t) {
			lds_float* p = p0 + 4 * (((t >> 2) & 1) * Box32<W>::si + ((t >> 1) & 1) * Box32<W>::sj + (t & 1));
			e[t][0] = p[0];
			e[t][1] = p[1];
			e[t][2] = p[2];
		}
	} else {
		const int sj = NODE ? g.N[2] + 1 : g.N[2], si = (NODE ? g.N[1] + 1 : g.N[1]) * sj;
		const int base = a * si + c * sj + d;
#pragma unroll
		for (int t = 0; t < 8; ++t) {
			const float4 v = f[base + ((t >> 2) & 1) * si + ((t >> 1) & 1) * sj + (t & 1)];
			e[t][0] = v.x;
			e[t][1] = v.y;
			e[t][2] = v.z;
		}
	}
	sum8(e, w0, w1, w2, out);
}

// The FP32 kernels take reciprocals and inverse square roots from the hardware's
// v_rcp_f32 / v_rsq_f32 (about 1 ulp) instead of correctly rounded divisions and square
// roots, which cost about ten instructions each.
__device__ inline float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }

// ------------------------------------------------------------------ coefficient boxes
// The CIC gather of a 2 x 2 x 2 stencil with corner values p[4i + 2j + k] at local
// coordinates (x, y, z) in [0, 1) (the weight of corner 1 along an axis) is the trilinear
// interpolant
//   F = (c0 + z cz) + y (cy + z cyz) + x ((cx + z cxz) + y (cxy + z cxyz)),
// seven fused multiply-adds per component once the 8 coefficients are formed, against 12
// weight products and 24 multiply-adds for the weighted corner sum of 3 components. The
// node map's local coordinate is the offset t in the cell, the cell map's t + 1/2 from the
// cell below (t < 1/2) or t - 1/2 from its own cell: the weights (1 - x, x) are the closed
// forms of correlationBspline (PIC++/correlation.cpp:824-835).
//
// A push workgroup whose particles lie in one (i, j) row of at most kCoefK cells stages, in
// LDS, the coefficients of every stencil origin that its gathers can use: E^n at the row's
// own nodes, E^{n+theta} with one node and B with two cells of margin (Picard iterates up
// to one cell away). A gather whose origin lies in its box reads 24 floats; any other forms
// the coefficients from the 8 corner values in HBM with the same operations, so both paths
// give the same bits.
constexpr int kCoefK = 12;

// p: the 8 corner values (x, y, z per point); c[2r], c[2r + 1] = component r's
// (c0, cz, cy, cyz), (cx, cxz, cxy, cxyz)
__device__ inline void trilinear_coefs(const float4 p[8], float4 c[6]) {
#pragma unroll
	for (int r = 0; r < 3; ++r) {
		float q[8];
#pragma unroll
		for (int t = 0; t < 8; ++t) q[t] = r == 0 ? p[t].x : r == 1 ? p[t].y : p[t].z;
		const float cz = q[1] - q[0], cy = q[2] - q[0], cx = q[4] - q[0];
		const float cyz = (q[3] - q[2]) - cz;
		const float cxz = (q[5] - q[4]) - cz;
		const float cxy = (q[6] - q[4]) - cy;
		const float cxyz = ((q[7] - q[6]) - (q[5] - q[4])) - cyz;
		c[2 * r] = make_float4(q[0], cz, cy, cyz);
		c[2 * r + 1] = make_float4(cx, cxz, cxy, cxyz);
	}
}

__device__ inline float trilinear_eval(const float a[4], const float b[4], float x, float y, float z) {
	const float g = fmaf(y, fmaf(z, a[3], a[2]), fmaf(z, a[1], a[0]));
	const float h = fmaf(y, fmaf(z, b[3], b[2]), fmaf(z, b[1], b[0]));
	return fmaf(x, h, g);
}

// node (nb, tn) and cell (cb, tc) map bases and local coordinates of a point at offset f
// from the lower corner of cell c (f may leave [0, 1) during the Picard iterations)
struct Loc32 {
	int nb, cb;
	float tn, tc;
};

__device__ inline void locs32(const int c[3], const float f[3], Loc32 L[3]) {
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		const float fl = floorf(f[a]);
		const int cc = c[a] + (int)fl;
		const float t = f[a] - fl;
		const bool lo = t < 0.5f;
		L[a].nb = cc;
		L[a].tn = t;
		L[a].cb = lo ? cc - 1 : cc;
		L[a].tc = lo ? t + 0.5f : t - 0.5f;
	}
}

// (W - 1) x (W - 1) x (K + W - 2) stencil origins of a W-point box, 24 floats each
template <int W>
struct CBox32 {
	int o[3];    // ghosted index of origin (0, 0, 0)
	int limk;    // last origin along k: K + W - 3
	int sj, si;  // origin strides along j and i
	int off;     // first float of the box in LDS
};

template <int W>
__device__ inline CBox32<W> make_cbox(int i0, int j0, int k0, int K, int off) {
	CBox32<W> b;
	constexpr int margin = W == 2 ? 0 : W - 3;
	b.o[0] = i0 - margin;
	b.o[1] = j0 - margin;
	b.o[2] = k0 - margin;
	b.limk = K + W - 3;
	b.sj = K + W - 2;
	b.si = (W - 1) * b.sj;
	b.off = off;
	return b;
}

// Gather with coefficients: from the box when the workgroup has one (use) and the origin
// lies in it, else from HBM with the bases clamped into the ghosted arrays (a clamp sets
// flag in why).
template <bool NODE, int W>
__device__ inline void gather_coef(const Grid& g, const float4* __restrict__ f, lds_float* lds, const CBox32<W>& b, bool use,
                                   const Loc32 L[3], int flag, int& why, float out[3]) {
	int a = NODE ? L[0].nb : L[0].cb, c = NODE ? L[1].nb : L[1].cb, d = NODE ? L[2].nb : L[2].cb;
	const float x = NODE ? L[0].tn : L[0].tc, y = NODE ? L[1].tn : L[1].tc, z = NODE ? L[2].tn : L[2].tc;
	const int la = a - b.o[0], lc = c - b.o[1], ld = d - b.o[2];
	if (use && (unsigned)la <= (unsigned)(W - 2) && (unsigned)lc <= (unsigned)(W - 2) && (unsigned)ld <= (unsigned)b.limk) {
		// one component's 8 coefficients at a time (fewer registers live)
		lds_float* p = lds + b.off + 24 * (la * b.si + lc * b.sj + ld);
#pragma unroll
		for (int r = 0; r < 3; ++r) {
			float c8[8];
#pragma unroll
			for (int t = 0; t < 8; ++t) c8[t] = p[8 * r + t];
			out[r] = trilinear_eval(c8, c8 + 4, x, y, z);
		}
		return;
	}
	const int lim0 = NODE ? g.N[0] - 1 : g.N[0] - 2, lim1 = NODE ? g.N[1] - 1 : g.N[1] - 2, lim2 = NODE ? g.N[2] - 1 : g.N[2] - 2;
	if (a < 0 || a > lim0 || c < 0 || c > lim1 || d < 0 || d > lim2) why |= flag;
	a = clampi(a, 0, lim0);
	c = clampi(c, 0, lim1);
	d = clampi(d, 0, lim2);
	const int sj = NODE ? g.N[2] + 1 : g.N[2], si = (NODE ? g.N[1] + 1 : g.N[1]) * sj;
	const int base = a * si + c * sj + d;
	float4 p[8], cf[6];
#pragma unroll
	for (int t = 0; t < 8; ++t) p[t] = f[base + ((t >> 2) & 1) * si + ((t >> 1) & 1) * sj + (t & 1)];
	trilinear_coefs(p, cf);
#pragma unroll
	for (int r = 0; r < 3; ++r) {
		const float c8[8] = {cf[2 * r].x, cf[2 * r].y, cf[2 * r].z, cf[2 * r].w,
		                     cf[2 * r + 1].x, cf[2 * r + 1].y, cf[2 * r + 1].z, cf[2 * r + 1].w};
		out[r] = trilinear_eval(c8, c8 + 4, x, y, z);
	}
}

// Stages the three coefficient boxes of a row of K cells from cell (i0, j0, k0): E^n (K
// origins), E^{n+theta} (9 (K + 2)) and B (16 (K + 3)), concatenated in LDS in origin order,
// so that origin o of the concatenation is at float 24 o. Corner indexes are clamped into
// the arrays (the boxes of a row of owned cells lie inside them).
__device__ inline void stage_coef_boxes(const Grid& g, int i0, int j0, int k0, int K, const float4* __restrict__ E,
                                        const float4* __restrict__ Et, const float4* __restrict__ B, float4* lds) {
	const int nEn = K, zt = K + 2, zb = K + 3;
	const int nEt = 9 * zt, total = nEn + nEt + 16 * zb;
	const float rzt = rcp32((float)zt), rzb = rcp32((float)zb);
	for (int o = threadIdx.x; o < total; o += blockDim.x) {
		// Origin o's box and local (la, lc, ld); row quotients from a float reciprocal, exact
		// for these sizes (the fraction of (e + 1/2) / len is at least 1 / (2 len) away from
		// an integer), and r / 3 = (11 r) >> 5 for r < 9.
		int a, c, d;
		bool node = true;
		const float4* src = E;
		if (o < nEn) {
			a = i0;
			c = j0;
			d = k0 + o;
		} else if (o < nEn + nEt) {
			const int e = o - nEn, r = (int)(((float)e + 0.5f) * rzt), la = (r * 11) >> 5;
			a = i0 - 1 + la;
			c = j0 - 1 + (r - 3 * la);
			d = k0 - 1 + (e - r * zt);
			src = Et;
		} else {
			const int e = o - nEn - nEt, r = (int)(((float)e + 0.5f) * rzb);
			a = i0 - 2 + (r >> 2);
			c = j0 - 2 + (r & 3);
			d = k0 - 2 + (e - r * zb);
			node = false;
			src = B;
		}
		const int l0 = node ? g.N[0] : g.N[0] - 1, l1 = node ? g.N[1] : g.N[1] - 1, l2 = node ? g.N[2] : g.N[2] - 1;
		const int sj = node ? g.N[2] + 1 : g.N[2], si = (node ? g.N[1] + 1 : g.N[1]) * sj;
		const int ia[2] = {clampi(a, 0, l0) * si, clampi(a + 1, 0, l0) * si};
		const int ic[2] = {clampi(c, 0, l1) * sj, clampi(c + 1, 0, l1) * sj};
		const int id[2] = {clampi(d, 0, l2), clampi(d + 1, 0, l2)};
		float4 p[8], cf[6];
#pragma unroll
		for (int t = 0; t < 8; ++t) p[t] = src[ia[(t >> 2) & 1] + ic[(t >> 1) & 1] + id[t & 1]];
		trilinear_coefs(p, cf);
#pragma unroll
		for (int r = 0; r < 6; ++r) lds[6 * o + r] = cf[r];
	}
}

// The 8 stencil values of a fixed stencil, loaded once and reused for every particle of
// a segment (all particles of one key share both stencil bases).
struct Stencil8 {
	float v[8][3];
};

__device__ inline void load_stencil(const float4* __restrict__ f, int base, int si, int sj, Stencil8& s) {
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		float4 e = f[base + ((t >> 2) & 1) * si + ((t >> 1) & 1) * sj + (t & 1)];
		s.v[t][0] = e.x;
		s.v[t][1] = e.y;
		s.v[t][2] = e.z;
	}
}

__device__ inline void interp8(const Stencil8& s, const float w0[2], const float w1[2], const float w2[2], float out[3]) {
	out[0] = out[1] = out[2] = 0.0f;
#pragma unroll
	for (int t = 0; t < 8; ++t) {
		float c = w0[(t >> 2) & 1] * w1[(t >> 1) & 1] * w2[t & 1];
		out[0] += s.v[t][0] * c;
		out[1] += s.v[t][1] * c;
		out[2] += s.v[t][2] * c;
	}
}

// alpha tensor (PIC++/mover.cpp:851-908) in FP32, with v and gamma from u = gamma v.
__device__ inline void alpha32(float beta, const float v[3], float gamma, const float E[3], const float B[3], float a[9]) {
	float G = beta * (E[0] * v[0] + E[1] * v[1] + E[2] * v[2]) + gamma;
	float iG = rcp32(G);
	float bs = beta * iG;
	float b2 = bs * bs;
	float inv = iG * rcp32(1.0f + b2 * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]));
	a[0] = (1.0f + b2 * B[0] * B[0]) * inv;
	a[1] = (b2 * B[0] * B[1] + bs * B[2]) * inv;
	a[2] = (b2 * B[0] * B[2] - bs * B[1]) * inv;
	a[3] = (b2 * B[1] * B[0] - bs * B[2]) * inv;
	a[4] = (1.0f + b2 * B[1] * B[1]) * inv;
	a[5] = (b2 * B[1] * B[2] + bs * B[0]) * inv;
	a[6] = (b2 * B[2] * B[0] + bs * B[1]) * inv;
	a[7] = (b2 * B[2] * B[1] - bs * B[0]) * inv;
	a[8] = (1.0f + b2 * B[2] * B[2]) * inv;
}

// Sums 8 per-lane values over each group of G consecutive lanes (8 <= G, G divides the
// workgroup) through LDS: group lane a < 8 gets the sum of value a over lanes 0, 1, ..., G-1,
// in that order, so the sums are deterministic. Every thread of the workgroup must call it;
// red holds 9 floats per thread (row stride 9 spreads the column reads over the LDS banks).
__device__ inline float group_sum8(const float v[8], float* red, int lt, int lane, int G) {
	__syncthreads();  // the previous chunk's reads are done
	float* row = red + lt * 9;
#pragma unroll
	for (int q = 0; q < 8; ++q) row[q] = v[q];
	__syncthreads();
	float s = 0.0f;
	if (lane < 8) {
		const float* col = red + (lt - lane) * 9 + lane;
		for (int l = 0; l < G; ++l) s += col[l * 9];
	}
	return s;
}

__device__ inline void load32(const ParticleArrays32& P, int i, float f[3], float u[3]) {
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		f[a] = P.off[a][i];
		u[a] = P.u[a][i];
	}
}

constexpr int kMaxTile = 16;  // cells per deposit workgroup

// ------------------------------------------------------------------ moments at x^n
// Deposit tiles: a workgroup serves T cells b = (i, j, k0 .. k0+T-1) of one z row, with a
// group of G lanes per cell. Lanes of a group take the cell's particles of each species
// round-robin (lane l: l, l+G, l+2G, ...), so they stay evenly loaded whatever the octant
// counts. Each lane sums its particles in that fixed order and the group reduces through LDS
// in a fixed lane order (group_sum8): the sums are deterministic.



// Node moments (J, mu) per cell. The E stencil (the cell's 8 nodes) is shared by all its
// particles and lives in registers; the B stencil depends on the octant and is read from
// the tile's 3 x 3 x (T+2) box of cells in LDS.
__global__ void __launch_bounds__(256) k_node_moments32(Grid g, SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ start,
                                                        const float4* __restrict__ E4, const float4* __restrict__ B4, float dt,
                                                        float theta, int G, int T, float* __restrict__ out) {
	extern __shared__ float4 Bn[];  // 3 x 3 x (T+2) cells around the tile
	__shared__ float red[256 * 9];
	const DualTile tile = dual_tile(g, T);
	const int lt = threadIdx.x;
	const int nk = T + 2;
	const long long c_first = cell_id(g, tile.i, tile.j, tile.k0);
	for (int q = lt; q < 9 * nk; q += blockDim.x) {
		int dk = q % nk, dj = (q / nk) % 3, di = q / (3 * nk);
		int ci = clampi(tile.i - 1 + di, 0, g.N[0] - 1), cj = clampi(tile.j - 1 + dj, 0, g.N[1] - 1);
		int ck = clampi(tile.k0 - 1 + dk, 0, g.N[2] - 1);
		Bn[q] = B4[cell_id(g, ci, cj, ck)];
	}
	__syncthreads();

	const int kb = lt / G, lane = lt % G;
	const int bk = tile.k0 + kb;
	const bool cell_ok = kb < T && bk < g.N[2];
	float acc[8][12];
#pragma unroll
	for (int a = 0; a < 8; ++a)
#pragma unroll
		for (int q = 0; q < 12; ++q) acc[a][q] = 0.0f;
	if (cell_ok) {
		const long long c = c_first + kb;
		const float inv_v = (float)(1.0 / g.volume);
		Stencil8 SE;
		{
			const int nj = g.N[2] + 1, ni = (g.N[1] + 1) * nj;
			load_stencil(E4, tile.i * ni + tile.j * nj + bk, ni, nj, SE);
		}
		for (int s = 0; s < sp.count; ++s) {
			const int lo = start[((long long)s * g.ncells + c) << 3], hi = start[((long long)s * g.ncells + c + 1) << 3];
			const float m = (float)sp.mass[s], q = (float)sp.charge[s], w = (float)sp.weight[s];
			const float beta = 0.5f * q * dt / m;
			const float omega = w * theta * dt * dt * 2.0f * kPiF * q * q / m * inv_v;
			const float pc = q * w * inv_v;
			for (int i = lo + lane; i < hi; i += G) {
				float f[3], u[3];
				load32(P, i, f, u);
				int bit[3];
				float nw[3][2], cw[3][2];
#pragma unroll
				for (int a3 = 0; a3 < 3; ++a3) {
					bit[a3] = f[a3] >= 0.5f;
					nw[a3][0] = 1.0f - f[a3];
					nw[a3][1] = f[a3];
					cw[a3][0] = (bit[a3] ? 1.5f : 0.5f) - f[a3];
					cw[a3][1] = f[a3] - (bit[a3] ? 0.5f : -0.5f);
				}
				float Ef[3], Bf[3] = {0.0f, 0.0f, 0.0f}, a[9];
				interp8(SE, nw[0], nw[1], nw[2], Ef);
				const int bb = (bit[0] * 3 + bit[1]) * nk + kb + bit[2];
#pragma unroll
				for (int t = 0; t < 8; ++t) {
					const int tx = (t >> 2) & 1, ty = (t >> 1) & 1, tz = t & 1;
					float4 bv = Bn[bb + (tx * 3 + ty) * nk + tz];
					float wt = cw[0][tx] * cw[1][ty] * cw[2][tz];
					Bf[0] += bv.x * wt;
					Bf[1] += bv.y * wt;
					Bf[2] += bv.z * wt;
				}
				const float g2 = 1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
				float ig = rsq32(g2);
				float gm = g2 * ig;
				float v[3] = {u[0] * ig, u[1] * ig, u[2] * ig};
				alpha32(beta, v, gm, Ef, Bf, a);
				float rv[3];
#pragma unroll
				for (int r = 0; r < 3; ++r) rv[r] = a[r * 3 + 0] * u[0] + a[r * 3 + 1] * u[1] + a[r * 3 + 2] * u[2];
#pragma unroll
				for (int ci = 0; ci < 2; ++ci)
#pragma unroll
					for (int cj = 0; cj < 2; ++cj)
#pragma unroll
						for (int ck = 0; ck < 2; ++ck) {
							int corner = ci * 4 + cj * 2 + ck;
							float corr = nw[0][ci] * nw[1][cj] * nw[2][ck];
							float fj = pc * corr;
#pragma unroll
							for (int r = 0; r < 3; ++r) acc[corner][r] += rv[r] * fj;
							float o = omega * corr;
#pragma unroll
							for (int r = 0; r < 9; ++r) acc[corner][3 + r] -= a[r] * o;
						}
			}
		}
	}
	const long long c = c_first + kb;
#pragma unroll
	for (int q = 0; q < 12; ++q) {
		float v8[8];
#pragma unroll
		for (int a = 0; a < 8; ++a) v8[a] = acc[a][q];
		const float sum = group_sum8(v8, red, lt, lane, G);
		if (cell_ok && lane < 8) out[((long long)q * 8 + lane) * g.ncells + c] = sum;
	}
}

// Cell moments (rhohat corners and the reference's pressure-tensor partials) per dual cell.
// The B stencil (cells b + t) is shared by the dual cell's particles and lives in registers;
// the E stencil depends on the source cell and is read from the tile's 3 x (Tj+2) x (Tk+2)
// box of nodes in LDS.
// DENS: the same pass also deposits the post-push density (updateDensityParameters, the
// planes of k_density32) into dens_out. Right after a push the particles sit at x^{n+1} and
// E^{n+1}, B^{n+1} exist, so one pass gives both this step's density and the next step's cell
// moments; rho and rhohat's corner sums are the same sums.
template <bool DENS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DENS ? 2 : 3))) k_cell_moments32(Grid g, SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ start,
                                                        const float4* __restrict__ E4, const float4* __restrict__ B4, float dt,
                                                        int G, int Tj, int Tk, float* __restrict__ rho_out,
                                                        float* __restrict__ pi_out, float* __restrict__ dens_out) {
	extern __shared__ float4 En[];  // 3 x (Tj+2) x (Tk+2) nodes from the tile's base corner
	__shared__ float red[256 * 9];
	const PlaneTile tile = plane_tile(g, Tj, Tk);
	const int lt = threadIdx.x;
	const int T = Tj * Tk, mj = Tj + 2, mk = Tk + 2;
	for (int q = lt; q < 3 * mj * mk; q += blockDim.x) {
		int dk = q % mk, dj = (q / mk) % mj, di = q / (mj * mk);
		int ni = clampi(tile.i + di, 0, g.N[0]), nj = clampi(tile.j0 + dj, 0, g.N[1]), nkk = clampi(tile.k0 + dk, 0, g.N[2]);
		En[q] = E4[node_id(g, ni, nj, nkk)];
	}
	__syncthreads();

	const int kb = lt / G, lane = lt % G;
	const int jb = kb / Tk, kk = kb % Tk;
	const int bj = tile.j0 + jb, bk = tile.k0 + kk;
	const bool base_ok = tile.ok && kb < T && bj < g.N[1] && bk < g.N[2];
	float ar[8], ap[8][6];
#pragma unroll
	for (int q = 0; q < 8; ++q) {
		ar[q] = 0.0f;
#pragma unroll
		for (int r = 0; r < 6; ++r) ap[q][r] = 0.0f;
	}
	const long long b = base_ok ? cell_id(g, tile.i, bj, bk) : 0;
	const float inv_v = (float)(1.0 / g.volume);
	Stencil8 SB;
	bool have_sb = false;
	for (int s = 0; s < sp.count; ++s) {  // every thread runs every species: DENS reduces per species
		float conc[8], en[8], bulk[8][3];
		if (DENS) {
#pragma unroll
			for (int q = 0; q < 8; ++q) conc[q] = en[q] = bulk[q][0] = bulk[q][1] = bulk[q][2] = 0.0f;
		}
		DualSegments d;
		if (base_ok) dual_segments(g, start, tile.i, bj, bk, s, d);
		if (base_ok && d.p[8] > 0) {
			if (!have_sb) {  // cells b .. b+1 exist whenever b has particles
				const int cj = g.N[2], ci = g.N[1] * cj;
				load_stencil(B4, tile.i * ci + bj * cj + bk, ci, cj, SB);
				have_sb = true;
			}
			const float m = (float)sp.mass[s], q = (float)sp.charge[s], w = (float)sp.weight[s];
			const float beta = 0.5f * q * dt / m;
			const float wv = w * inv_v;
			const float pc = q * wv;
			// The fused pass (DENS, 2 waves per SIMD) is software pipelined: the next particle's
			// loads are issued before this one's work (a lane past the end reloads a valid particle
			// and drops it). The plain pass keeps 3 waves only without the extra registers.
			const int nv = d.p[8];
			auto fetch = [&](int v, float f[3], float u[3], int& a) {
				v = v < nv ? v : nv - 1;
				int off = d.off[0];
				a = 0;
#pragma unroll
				for (int qq = 1; qq < 8; ++qq)
					if (v >= d.p[qq]) {
						a = qq;
						off = d.off[qq];
					}
				load32(P, v + off, f, u);
			};
			float fq[3], uq[3];
			int aq = 0;
			if constexpr (DENS) fetch(lane, fq, uq, aq);
			for (int v = lane; v < nv; v += G) {
				float f[3], u[3];
				int a;
				if constexpr (DENS) {
					a = aq;
#pragma unroll
					for (int c3 = 0; c3 < 3; ++c3) {
						f[c3] = fq[c3];
						u[c3] = uq[c3];
					}
					fetch(v + G, fq, uq, aq);
					__builtin_amdgcn_sched_barrier(0);  // issue those loads here
				} else {
					fetch(v, f, u, a);
				}
				const int ax = (a >> 2) & 1, ay = (a >> 1) & 1, az = a & 1;
				const int bit[3] = {1 - ax, 1 - ay, 1 - az};
				float nw[3][2], cw[3][2];
#pragma unroll
				for (int a3 = 0; a3 < 3; ++a3) {
					nw[a3][0] = 1.0f - f[a3];
					nw[a3][1] = f[a3];
					cw[a3][0] = (bit[a3] ? 1.5f : 0.5f) - f[a3];
					cw[a3][1] = f[a3] - (bit[a3] ? 0.5f : -0.5f);
				}
				float Ef[3] = {0.0f, 0.0f, 0.0f}, Bf[3], al[9];
				const int eb = (ax * mj + jb + ay) * mk + kk + az;
#pragma unroll
				for (int t = 0; t < 8; ++t) {
					const int tx = (t >> 2) & 1, ty = (t >> 1) & 1, tz = t & 1;
					float4 ev = En[eb + (tx * mj + ty) * mk + tz];
					float wt = nw[0][tx] * nw[1][ty] * nw[2][tz];
					Ef[0] += ev.x * wt;
					Ef[1] += ev.y * wt;
					Ef[2] += ev.z * wt;
				}
				interp8(SB, cw[0], cw[1], cw[2], Bf);
				const float g2 = 1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
				float ig = rsq32(g2);
				float gm = g2 * ig;
				float vv[3] = {u[0] * ig, u[1] * ig, u[2] * ig};
				alpha32(beta, vv, gm, Ef, Bf, al);
				float rv[3];
#pragma unroll
				for (int r = 0; r < 3; ++r) rv[r] = al[r * 3 + 0] * u[0] + al[r * 3 + 1] * u[1] + al[r * 3 + 2] * u[2];
				float Tt[6] = {rv[0] * rv[0], rv[0] * rv[1], rv[0] * rv[2], rv[1] * rv[1], rv[1] * rv[2], rv[2] * rv[2]};
#pragma unroll
				for (int ti = 0; ti < 2; ++ti)
#pragma unroll
					for (int tj = 0; tj < 2; ++tj)
#pragma unroll
						for (int tk = 0; tk < 2; ++tk) {
							int tt = ti * 4 + tj * 2 + tk;
							const float corr = cw[0][ti] * cw[1][tj] * cw[2][tk];
							const float fr = pc * corr;
							ar[tt] += fr;
#pragma unroll
							for (int r = 0; r < 6; ++r) ap[tt][r] += fr * Tt[r];
							if (DENS) {
								const float c = wv * corr;
								conc[tt] += c;
								en[tt] += c * gm * m;
#pragma unroll
								for (int r = 0; r < 3; ++r) bulk[tt][r] += c * u[r] * m;
							}
						}
				// the next particle's values are needed from here on only: the loads stay ahead of
				// this particle's work instead of being sunk into the next iteration
				if constexpr (DENS) asm volatile("" : "+v"(fq[0]), "+v"(fq[1]), "+v"(fq[2]), "+v"(uq[0]), "+v"(uq[1]), "+v"(uq[2]));
			}
		}
		if (DENS) {  // planes per species [conc, energy, bulk x3] x 8 corners, as k_density32
			const long long base = (long long)s * 5 * 8;
			{
				const float sum = group_sum8(conc, red, lt, lane, G);
				if (base_ok && lane < 8) dens_out[(base + 0 * 8 + lane) * g.ncells + b] = sum;
			}
			{
				const float sum = group_sum8(en, red, lt, lane, G);
				if (base_ok && lane < 8) dens_out[(base + 1 * 8 + lane) * g.ncells + b] = sum;
			}
#pragma unroll
			for (int r = 0; r < 3; ++r) {
				float v8[8];
#pragma unroll
				for (int q = 0; q < 8; ++q) v8[q] = bulk[q][r];
				const float sum = group_sum8(v8, red, lt, lane, G);
				if (base_ok && lane < 8) dens_out[(base + (2 + r) * 8 + lane) * g.ncells + b] = sum;
			}
		}
	}
	{
		const float sum = group_sum8(ar, red, lt, lane, G);
		if (base_ok && lane < 8) {
			rho_out[(long long)lane * g.ncells + b] = sum;
			if (DENS) dens_out[((long long)sp.count * 5 * 8 + lane) * g.ncells + b] = sum;
		}
	}
#pragma unroll
	for (int r = 0; r < 6; ++r) {
		float v8[8];
#pragma unroll
		for (int q = 0; q < 8; ++q) v8[q] = ap[q][r];
		const float sum = group_sum8(v8, red, lt, lane, G);
		if (base_ok && lane < 8) pi_out[((long long)lane * 6 + r) * g.ncells + b] = sum;
	}
}

// Post-push density (updateDensityParameters) per dual cell; lanes of a group share each
// species' particles round-robin, so the per-species sums reduce over the whole group.
// Output planes (float): per species [conc, energy, bulk x3] x 8 corners, then rho x 8.
__global__ void __launch_bounds__(256) k_density32(Grid g, SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ start, int G,
                                                   int Tj, int Tk, float* __restrict__ out) {
	__shared__ float red[256 * 9];
	const PlaneTile tile = plane_tile(g, Tj, Tk);
	const int lt = threadIdx.x;
	const int T = Tj * Tk;
	const int kb = lt / G, lane = lt % G;
	const int bj = tile.j0 + kb / Tk, bk = tile.k0 + kb % Tk;
	const bool base_ok = tile.ok && kb < T && bj < g.N[1] && bk < g.N[2];
	const long long b = base_ok ? cell_id(g, tile.i, bj, bk) : 0;
	const float wvol = (float)(1.0 / g.volume);
	float rho[8];
#pragma unroll
	for (int q = 0; q < 8; ++q) rho[q] = 0.0f;
	for (int s = 0; s < sp.count; ++s) {
		float conc[8], en[8], bulk[8][3];
#pragma unroll
		for (int q = 0; q < 8; ++q) {
			conc[q] = en[q] = 0.0f;
			bulk[q][0] = bulk[q][1] = bulk[q][2] = 0.0f;
		}
		if (base_ok) {
			DualSegments d;
			dual_segments(g, start, tile.i, bj, bk, s, d);
			const float m = (float)sp.mass[s];
			const float wv = (float)sp.weight[s] * wvol;
			const float pc = (float)sp.charge[s] * wv;
			for (int v = lane; v < d.p[8]; v += G) {
				int a = 0, off = d.off[0];
#pragma unroll
				for (int qq = 1; qq < 8; ++qq)
					if (v >= d.p[qq]) {
						a = qq;
						off = d.off[qq];
					}
				float f[3], u[3];
				load32(P, v + off, f, u);
				const int bit[3] = {1 - ((a >> 2) & 1), 1 - ((a >> 1) & 1), 1 - (a & 1)};
				float cw[3][2];
#pragma unroll
				for (int a3 = 0; a3 < 3; ++a3) {
					cw[a3][0] = (bit[a3] ? 1.5f : 0.5f) - f[a3];
					cw[a3][1] = f[a3] - (bit[a3] ? 0.5f : -0.5f);
				}
				const float g2 = 1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
				float gm = g2 * rsq32(g2);
#pragma unroll
				for (int ti = 0; ti < 2; ++ti)
#pragma unroll
					for (int tj = 0; tj < 2; ++tj)
#pragma unroll
						for (int tk = 0; tk < 2; ++tk) {
							int tt = ti * 4 + tj * 2 + tk;
							float corr = cw[0][ti] * cw[1][tj] * cw[2][tk];
							float c = wv * corr;
							conc[tt] += c;
							en[tt] += c * gm * m;
							rho[tt] += pc * corr;
#pragma unroll
							for (int r = 0; r < 3; ++r) bulk[tt][r] += c * u[r] * m;
						}
			}
		}
		const long long base = (long long)s * 5 * 8;
		{
			const float sum = group_sum8(conc, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + 0 * 8 + lane) * g.ncells + b] = sum;
		}
		{
			const float sum = group_sum8(en, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + 1 * 8 + lane) * g.ncells + b] = sum;
		}
#pragma unroll
		for (int r = 0; r < 3; ++r) {
			float v8[8];
#pragma unroll
			for (int q = 0; q < 8; ++q) v8[q] = bulk[q][r];
			const float sum = group_sum8(v8, red, lt, lane, G);
			if (base_ok && lane < 8) out[(base + (2 + r) * 8 + lane) * g.ncells + b] = sum;
		}
	}
	{
		const float sum = group_sum8(rho, red, lt, lane, G);
		if (base_ok && lane < 8) out[((long long)sp.count * 5 * 8 + lane) * g.ncells + b] = sum;
	}
}

// ------------------------------------------------------------------ mover
// Round-up multiplier division (Granlund and Montgomery, PLDI 1994, fig. 4.1): n / d for any
// 32-bit n, with l = ceil(log2 d) >= 1 and m = floor(2^32 (2^l - d) / d) + 1.
__device__ inline uint32_t div_magic(uint32_t n, uint32_t m, int l) {
	const uint32_t t = __umulhi(n, m);
	return (t + ((n - t) >> 1)) >> (l - 1);
}

__device__ inline void cell_ijk_magic(const Grid& g, const Push32Consts& K, int cell, int c[3]) {
	const uint32_t q1 = div_magic((uint32_t)cell, K.m2, K.l2);  // cell / N[2]
	const uint32_t q2 = div_magic(q1, K.m1, K.l1);              // cell / (N[2] N[1])
	c[2] = cell - (int)q1 * g.N[2];
	c[1] = (int)q1 - (int)q2 * g.N[1];
	c[0] = (int)q2;
}

// moveParticle (PIC++/mover.cpp:194-491) on cell offsets, one thread per particle; alpha is
// recomputed from E^n, B^n at x^n. Momentum-space constants: du = (q dt / m) (E + v x B).
template <bool WALL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WALL ? 1 : 4))) k_push32(Grid g, SpeciesTable sp, Push32Consts K, ParticleArrays32 P,
                                                const int* __restrict__ species_start, const float4* __restrict__ E,
                                                const float4* __restrict__ Et, const float4* __restrict__ B, float dt, float eta,
                                                uint32_t* __restrict__ keys, uint64_t* __restrict__ mask, int* __restrict__ status,
                                                int* __restrict__ dead, uint32_t* __restrict__ esc,
                                                uint64_t* __restrict__ bounce_mask) {
	constexpr int oEt = Box32<2>::size, oB = Box32<2>::size + Box32<4>::size;
	constexpr int kPointF4 = Box32<2>::size + Box32<4>::size + Box32<5>::size;
	constexpr int kCoefF4 = 6 * (kCoefK + 9 * (kCoefK + 2) + 16 * (kCoefK + 3));
	// E^n, E^{n+theta}, B: point boxes or coefficient boxes
	__shared__ float4 boxes[kPointF4 > kCoefF4 ? kPointF4 : kCoefF4];
	__shared__ int row_ends[2];
	const long long b0 = (long long)blockIdx.x * blockDim.x;
	const long long idx = b0 + threadIdx.x;
	const int last = (int)min((long long)blockDim.x - 1, P.count - 1 - b0);
	if ((int)threadIdx.x == 0) row_ends[0] = P.cell[b0];
	if ((int)threadIdx.x == last) row_ends[1] = P.cell[b0 + last];
	__syncthreads();
	int rf[3], rl[3];
	cell_ijk_magic(g, K, row_ends[0], rf);
	cell_ijk_magic(g, K, row_ends[1], rl);
	const int KB = rl[2] - rf[2] + 1;
	// Rows of at most kCoefK cells gather with coefficient boxes, rows of up to kBoxK cells
	// with point boxes (the coefficient boxes of a long row would not fit the LDS budget of
	// four workgroups per CU), any other workgroup from HBM.
	const bool one_row = rf[0] == rl[0] && rf[1] == rl[1] && KB >= 1;
	const bool coef = one_row && KB <= kCoefK;
	const bool use_box = one_row && !coef && KB <= kBoxK;
	const Box32<2> bEn = make_box<2>(rf[0], rf[1], rf[2], KB);
	const Box32<4> bEt = make_box<4>(rf[0], rf[1], rf[2], KB);
	const Box32<5> bB = make_box<5>(rf[0], rf[1], rf[2], KB);
	const CBox32<2> cEn = make_cbox<2>(rf[0], rf[1], rf[2], KB, 0);
	const CBox32<4> cEt = make_cbox<4>(rf[0], rf[1], rf[2], KB, 24 * KB);
	const CBox32<5> cB = make_cbox<5>(rf[0], rf[1], rf[2], KB, 24 * (KB + 9 * (KB + 2)));
	if (coef) {
		stage_coef_boxes(g, rf[0], rf[1], rf[2], KB, E, Et, B, boxes);
	} else if (use_box) {
		// Stage the boxes: 4, 16 and 25 rows along k of K + W - 1 points (W = 2, 4, 5). Thread t
		// loads point t of each box in the same pass, so its three loads are in flight together.
		// Point e of a box is row r = e / len, point ld = e - r len (the quotient from a float
		// reciprocal, exact for these sizes: the fraction of (e + 1/2) / len is at least
		// 1 / (2 len) away from an integer); a row r of a box of width W is (la, lc) = (r / W,
		// r % W), r / 5 = (13 r) >> 6 for r < 25. Ghosted indexes are clamped into the arrays.
		const int len0 = KB + 1, len1 = KB + 3, len2 = KB + 4;
		const int n0 = 4 * len0, n1 = 16 * len1, n2 = 25 * len2;
		const float rl0 = rcp32((float)len0), rl1 = rcp32((float)len1), rl2 = rcp32((float)len2);
		const int nj = g.N[1] + 1, nk = g.N[2] + 1;  // node strides (cells: N[1], N[2])
		for (int e = threadIdx.x; e < n2; e += blockDim.x) {  // n2 is the largest
			float4 v0, v1, v2;
			int o0 = 0, o1 = 0, o2 = 0;
			if (e < n0) {  // E^n on the row's nodes
				const int r = (int)(((float)e + 0.5f) * rl0), ld = e - r * len0;
				const int a = clampi(rf[0] + (r >> 1), 0, g.N[0]), c = clampi(rf[1] + (r & 1), 0, g.N[1]),
				          d = clampi(rf[2] + ld, 0, g.N[2]);
				v0 = E[(a * nj + c) * nk + d];
				o0 = r * Box32<2>::sj + ld;
			}
			if (e < n1) {  // E^{n+theta}, one node of margin
				const int r = (int)(((float)e + 0.5f) * rl1), ld = e - r * len1;
				const int a = clampi(rf[0] - 1 + (r >> 2), 0, g.N[0]), c = clampi(rf[1] - 1 + (r & 3), 0, g.N[1]),
				          d = clampi(rf[2] - 1 + ld, 0, g.N[2]);
				v1 = Et[(a * nj + c) * nk + d];
				o1 = oEt + r * Box32<4>::sj + ld;
			}
			{  // B, two cells of margin
				const int r = (int)(((float)e + 0.5f) * rl2), ld = e - r * len2;
				const int la = (r * 13) >> 6, lc = r - 5 * la;
				const int a = clampi(rf[0] - 2 + la, 0, g.N[0] - 1), c = clampi(rf[1] - 2 + lc, 0, g.N[1] - 1),
				          d = clampi(rf[2] - 2 + ld, 0, g.N[2] - 1);
				v2 = B[(a * g.N[1] + c) * g.N[2] + d];
				o2 = oB + r * Box32<5>::sj + ld;
			}
			if (e < n0) boxes[o0] = v0;
			if (e < n1) boxes[o1] = v1;
			boxes[o2] = v2;
		}
	}
	__syncthreads();
	if (idx >= P.count) return;
	lds_float* lEn = (lds_float*)boxes;
	lds_float* lEt = (lds_float*)(boxes + oEt);
	lds_float* lB = (lds_float*)(boxes + oB);
	lds_float* lco = (lds_float*)boxes;
	// the particle's push, with the coefficient-box (COEF) or point-box gathers; coef is the
	// same for the whole workgroup
	auto body = [&](auto coef_tag) {
	constexpr bool COEF = decltype(coef_tag)::value;
	int i = (int)idx;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	const float qdm = K.qdm[s], beta = K.beta[s];
	const float inv_d[3] = {K.inv_d[0], K.inv_d[1], K.inv_d[2]};
	const int cell0 = P.cell[i];
	int cc[3];
	cell_ijk_magic(g, K, cell0, cc);
	float f[3], u[3];
	load32(P, i, f, u);
	const uint32_t key_s = (uint32_t)s * (uint32_t)g.ncells;  // keys fit 32 bits (checked on the host)
	const uint32_t key0 = ((key_s + (uint32_t)cell0) << 3) | ((f[0] >= 0.5f) << 2) | ((f[1] >= 0.5f) << 1) | (f[2] >= 0.5f);
	Map32 mp[3];
	Loc32 L[3];
	int why = 0;  // failed checks: 1 start stencil, 2 Picard stencil, 4 << c new cell
	float En[3], Bf[3], a[9], Ef[3];
	if constexpr (COEF) {
		locs32(cc, f, L);
		gather_coef<true, 2>(g, E, lco, cEn, true, L, 1, why, En);
		gather_coef<false, 5>(g, B, lco, cB, true, L, 1, why, Bf);
		gather_coef<true, 4>(g, Et, lco, cEt, true, L, 1, why, Ef);
	} else {
		if (!maps32(g, cc, f, mp)) why = 1;
		gather_boxed<true, 2>(g, E, lEn, bEn, use_box, mp, En);
		gather_boxed<false, 5>(g, B, lB, bB, use_box, mp, Bf);
		gather_boxed<true, 4>(g, Et, lEt, bEt, use_box, mp, Ef);
	}
	float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
	float ig = rsq32(1.0f + u2);
	float gm = (1.0f + u2) * ig;
	float v[3] = {u[0] * ig, u[1] * ig, u[2] * ig};
	alpha32(beta, v, gm, En, Bf, a);
	// predictor (mover.cpp:234-245), with Particle::evaluateVelocity's nonrelativistic branch
	float nu[3];
	nu[0] = u[0] + (Ef[0] + (v[1] * Bf[2] - v[2] * Bf[1])) * qdm;
	nu[1] = u[1] + (Ef[1] + (-v[0] * Bf[2] + v[2] * Bf[0])) * qdm;
	nu[2] = u[2] + (Ef[2] + (v[0] * Bf[1] - v[1] * Bf[0])) * qdm;
	float nu2 = nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2];
	float nig = nu2 < 0.0001f ? 1.0f : rsq32(1.0f + nu2);
	const float nup[3] = {nu[0], nu[1], nu[2]};  // the predictor's u (the Picard loop overwrites nu)
	float mid[3], xh[3];
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		mid[c] = v[c] * (1.0f - eta) + nu[c] * nig * eta;
		xh[c] = f[c] + mid[c] * (eta * dt) * inv_d[c];
	}
	// Wall boundary (x node 2): a predictor or Picard position behind the wall ends the push
	// there, reflected with |v_x| of the predictor velocity and the rest of the step
	// (PIC++/mover.cpp:248-258, :315-331, Particle::setMomentumByV). Positions are in cell units
	// relative to the start cell, so the wall sits at 2 - cc[0].
	const float rest = (1.0f - eta) * dt;
	float pos[3], unew[3];
	bool reflected = false;
	auto reflect_at_wall = [&](const float xc[3]) {
		pos[0] = (float)(4 - 2 * cc[0]) - xc[0] + fabsf(mid[0] * rest) * inv_d[0];
		pos[1] = xc[1] + mid[1] * rest * inv_d[1];
		pos[2] = xc[2] + mid[2] * rest * inv_d[2];
		// u = v / sqrt(1 - v^2) of the reflected predictor velocity v = nup nig. For v^2 >= 1e-4
		// that is nup itself (|v| is unchanged by the reflection and 1 - v^2 = nig^2), which
		// avoids forming 1 - v^2 in FP32 at high gamma; below, the nonrelativistic branch u = v.
		const float v2 = nu2 * nig * nig;
		const float sc = v2 < 0.0001f ? nig : 1.0f;
		unew[0] = fabsf(nup[0]) * sc;
		unew[1] = nup[1] * sc;
		unew[2] = nup[2] * sc;
		reflected = true;
	};
	if (WALL && xh[0] < (float)(2 - cc[0])) reflect_at_wall(xh);
	float vhat[3];
#pragma unroll
	for (int r = 0; r < 3; ++r) vhat[r] = a[r * 3 + 0] * u[0] + a[r * 3 + 1] * u[1] + a[r * 3 + 2] * u[2];
	// the stopping rule on squares: |du| > 1e-14 |u| and |dx| > 1e-8 dx (mover.cpp:292)
	float du[3] = {nu[0] - u[0], nu[1] - u[1], nu[2] - u[2]};
	float err2 = du[0] * du[0] + du[1] * du[1] + du[2] * du[2];
	float prevX[3] = {xh[0], xh[1], xh[2]};
	float cerr2 = 1.0f;  // in cell units (coordinatesError = deltaX)
	const float inv_d02 = inv_d[0] * inv_d[0];
	for (int it = 0; !reflected && it < 3 && err2 > 1e-28f * u2 && cerr2 > 1e-16f; ++it) {
		float prevU[3] = {nu[0], nu[1], nu[2]};
		float rE[3];
#pragma unroll
		for (int r = 0; r < 3; ++r) rE[r] = a[r * 3 + 0] * Ef[0] + a[r * 3 + 1] * Ef[1] + a[r * 3 + 2] * Ef[2];
		float nx[3];
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			mid[c] = vhat[c] + rE[c] * beta;
			nx[c] = f[c] + mid[c] * (eta * dt) * inv_d[c];
		}
		if (WALL && nx[0] < (float)(2 - cc[0])) {
			reflect_at_wall(nx);
			break;
		}
		if constexpr (COEF) {
			locs32(cc, nx, L);
			gather_coef<true, 4>(g, Et, lco, cEt, true, L, 2, why, Ef);
			gather_coef<false, 5>(g, B, lco, cB, true, L, 2, why, Bf);
		} else {
			if (!maps32(g, cc, nx, mp)) why |= 2;
			gather_boxed<true, 4>(g, Et, lEt, bEt, use_box, mp, Ef);
			gather_boxed<false, 5>(g, B, lB, bB, use_box, mp, Bf);
		}
		nu[0] = u[0] + (Ef[0] + (mid[1] * Bf[2] - mid[2] * Bf[1])) * qdm;
		nu[1] = u[1] + (Ef[1] + (-mid[0] * Bf[2] + mid[2] * Bf[0])) * qdm;
		nu[2] = u[2] + (Ef[2] + (mid[0] * Bf[1] - mid[1] * Bf[0])) * qdm;
		float d0 = prevU[0] - nu[0], d1 = prevU[1] - nu[1], d2 = prevU[2] - nu[2];
		err2 = d0 * d0 + d1 * d1 + d2 * d2;
		float e0 = (prevX[0] - nx[0]) * K.d[0], e1 = (prevX[1] - nx[1]) * K.d[1], e2 = (prevX[2] - nx[2]) * K.d[2];
		cerr2 = (e0 * e0 + e1 * e1 + e2 * e2) * inv_d02;
#pragma unroll
		for (int c = 0; c < 3; ++c) prevX[c] = nx[c];
	}
	if (!reflected) {
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			pos[c] = f[c] + mid[c] * dt * inv_d[c];
			unew[c] = nu[c];
		}
		if (WALL && pos[0] < (float)(2 - cc[0])) {  // final position behind the wall (mover.cpp:435-441)
			pos[0] = (float)(4 - 2 * cc[0]) - pos[0];
			unew[0] = -unew[0];
			reflected = true;
		}
	}
	// Out through the free right end: removed (mover.cpp:447-458). The reference removes x > node
	// n+2; a position exactly on that node is removed here too, so that every kept particle has
	// an owned cell.
	bool mv = false;  // the particle leaves its (species, cell, octant) segment
	if (WALL && pos[0] >= (float)(g.n[0] + 2 - cc[0])) {
		keys[i] = 0xFFFFFFFFu;
		mv = true;
		// listed for replacement, with its absolute y, z in cell units, and with its new momentum
		// (what leaves the box, for the theoretical energy)
		P.off[1][i] = (float)cc[1] + pos[1];
		P.off[2][i] = (float)cc[2] + pos[2];
#pragma unroll
		for (int c = 0; c < 3; ++c) P.u[c][i] = unew[c];
		esc[atomicAdd(dead, 1)] = (uint32_t)i;
	} else {
		// x^{n+1} = x^n + mid dt, carried into the integer cell; one-rank periodic wrap (x too
		// unless walled)
		int nc[3];
#pragma unroll
		for (int c = 0; c < 3; ++c) {
			float x = pos[c];
			float fl = floorf(x);
			x -= fl;
			if (x >= 1.0f) {  // rounding of x - floor(x) up to 1
				x = 0.0f;
				fl += 1.0f;
			}
			int ci = cc[c] + (int)fl;
			const bool walled = WALL && c == 0;
			if (!walled && ci < 2) ci += g.n[c];
			if (!walled && ci > g.n[c] + 1) ci -= g.n[c];
			if (ci < 2 || ci > g.n[c] + 1) why |= 4 << c;
			nc[c] = ci;
			f[c] = x;
		}
		if (why) {
			// A failed check ends the step with an error. The first failing particle keeps its
			// state from before the push, and status[2..4] hold the failed checks, its index and
			// species for the host's message.
			atomicOr(status, 8);
			if (atomicCAS(status + 1, 0, 1) == 0) {
				status[2] = why;
				status[3] = i;
				status[4] = s;
			}
		} else {
			const int cell = (nc[0] * g.N[1] + nc[1]) * g.N[2] + nc[2];
			P.cell[i] = cell;
#pragma unroll
			for (int c = 0; c < 3; ++c) {
				P.off[c][i] = f[c];
				P.u[c][i] = unew[c];
			}
			const uint32_t key = ((key_s + (uint32_t)cell) << 3) | ((f[0] >= 0.5f) << 2) | ((f[1] >= 0.5f) << 1) | (f[2] >= 0.5f);
			keys[i] = key;
			mv = key != key0;
		}
	}
	// one mover bit per particle: bit i & 63 of word i >> 6, written by the wave's first lane
	// (waves are 64 consecutive particles; lanes past the end contribute zeros)
	const unsigned long long bits = __ballot(mv);
	if ((threadIdx.x & 63) == 0) mask[i >> 6] = bits;
	if (WALL) {  // and one bit per particle reflected at the wall in this push
		const unsigned long long rb = __ballot(reflected);
		if ((threadIdx.x & 63) == 0) bounce_mask[i >> 6] = rb;
	}
	};
	if (coef)
		body(std::integral_constant<bool, true>());
	else
		body(std::integral_constant<bool, false>());
}

// ------------------------------------------------------------------ rebin after a push
// Few particles change segment in one step, so the cell-sorted order is restored without a
// full sort. New segment k holds its stayers in their old order, then its arrivals in the
// order of their old indices: deterministic, and the same order a stable sort by key gives
// inside each group. The push leaves one mover bit per particle (mask word i >> 6, bit
// i & 63); F(i), the number of movers before particle i, is the exclusive prefix count of the
// words' populations plus the population of the word's bits below i.

__device__ inline bool is_mover(const uint64_t* __restrict__ mask, long long i) { return (mask[i >> 6] >> (i & 63)) & 1ull; }

// F(i) for 0 <= i <= n; prefix has one entry per word plus the total
__device__ inline int movers_before(const uint64_t* __restrict__ mask, const int* __restrict__ prefix, long long i) {
	const long long w = i >> 6;
	const int b = (int)(i & 63);
	const int below = b ? __popcll(mask[w] & (~0ull >> (64 - b))) : 0;
	return prefix[w] + below;
}

// word populations; counts[nwords] = 0 closes the scan
__global__ void k_mask_counts(const uint64_t* __restrict__ mask, long long nwords, int* __restrict__ counts) {
	long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (w > nwords) return;
	counts[w] = w < nwords ? __popcll(mask[w]) : 0;
}

// sets the mover bits of particles lo .. hi-1 (appended after the push), keeping the bits
// below lo and clearing those from hi on in the last word; one thread per word
__global__ void k_mark_movers(uint64_t* __restrict__ mask, long long lo, long long hi) {
	const long long w = (lo >> 6) + (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (w > ((hi - 1) >> 6)) return;
	const long long a = w * 64 > lo ? w * 64 : lo, b = (w + 1) * 64 < hi ? (w + 1) * 64 : hi;
	const uint64_t set = (~0ull >> (64 - (int)(b - a))) << (int)(a - w * 64);
	const uint64_t keep = w * 64 < lo ? mask[w] & (~0ull >> (64 - (int)(lo - w * 64))) : 0ull;
	mask[w] = keep | set;
}

// movers in old-index order: mk[F(i)] = key, mi[F(i)] = i; removed particles (keys >= nkeys)
// take the key nkeys, which has no segment. One thread per mask word: a thread per particle
// would launch a million workgroups for a few percent of movers.
__global__ void k_gather_movers(const uint64_t* __restrict__ mask, const int* __restrict__ prefix, const uint32_t* __restrict__ keys,
                                long long n, long long nwords, uint32_t nkeys, uint32_t* __restrict__ mk, uint32_t* __restrict__ mi) {
	const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (w >= nwords) return;
	uint64_t bits = mask[w];
	int m = prefix[w];
	while (bits) {
		const long long i = w * 64 + (__ffsll((unsigned long long)bits) - 1);  // lowest set bit first
		bits &= bits - 1;
		if (i >= n) break;
		const uint32_t k = keys[i];
		mk[m] = k < nkeys ? k : nkeys;
		mi[m] = (uint32_t)i;
		++m;
	}
}

// first[k] = index of the first sorted mover with key k
__global__ void k_first_of_run(const uint32_t* __restrict__ mk, long long M, int* __restrict__ first) {
	long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= M) return;
	if (m == 0 || mk[m - 1] != mk[m]) first[mk[m]] = (int)m;
}

// counts[k] = old segment length minus its leavers, plus its arrivals; counts[nkeys] = 0
__global__ void k_rebin_counts(const int* __restrict__ start_old, const uint64_t* __restrict__ mask, const int* __restrict__ prefix,
                               const int* __restrict__ arrive, long long nkeys, int* __restrict__ counts) {
	long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k > nkeys) return;
	if (k == nkeys) {
		counts[k] = 0;
		return;
	}
	int lo = start_old[k], hi = start_old[k + 1];
	counts[k] = (hi - lo) - (movers_before(mask, prefix, hi) - movers_before(mask, prefix, lo)) + arrive[k];
}

// The destinations are written as a source index per new position (perm[to] = i), and the
// particles then move with one coalesced gather (k_permute32): scattering the 32-byte
// particles directly writes partial cache lines at every segment boundary. One thread per
// segment k writes its places from start_new[k] on: the stayers (old indexes start_old[k] ..
// start_old[k+1]-1 without a mover bit) in order, then the arrivals mi[first[k] ..
// first[k]+arrive[k]-1] (the movers sorted by new key, stable in old index). A workgroup's
// places form one range of perm; when it fits kPlaceStage entries it is assembled in LDS and
// written out coalesced (the per-thread writes are strided by the segment lengths), otherwise
// the threads write perm directly.
constexpr int kPlaceStage = 4096;
__global__ void __launch_bounds__(256) k_rebin_place(const int* __restrict__ start_old, const int* __restrict__ start_new,
                                                     const uint64_t* __restrict__ mask, const int* __restrict__ arrive,
                                                     const int* __restrict__ first, const uint32_t* __restrict__ mi, long long nkeys,
                                                     uint32_t* __restrict__ perm) {
	__shared__ uint32_t stage[kPlaceStage];
	const long long k0 = (long long)blockIdx.x * blockDim.x;
	const long long k = k0 + threadIdx.x;
	const long long kend = k0 + blockDim.x < nkeys ? k0 + blockDim.x : nkeys;
	const long long base = start_new[k0], top = start_new[kend];  // start_new has nkeys + 1 entries
	const bool staged = top - base <= kPlaceStage;  // the same for the whole workgroup
	auto put = [&](long long d, uint32_t v) {
		if (staged)
			stage[d - base] = v;
		else
			perm[d] = v;
	};
	if (k < nkeys) {
		const int lo = start_old[k], hi = start_old[k + 1];
		long long dst = start_new[k];
		for (long long w = lo >> 6; w * 64 < hi; ++w) {
			// the word's bits inside [lo, hi) that are not movers
			const int a = lo > w * 64 ? (int)(lo - w * 64) : 0, b = hi < (w + 1) * 64 ? (int)(hi - w * 64) : 64;
			uint64_t stay = ~mask[w] & ((b == 64 ? ~0ull : ((1ull << b) - 1)) & ~((1ull << a) - 1));
			while (stay) {
				put(dst++, (uint32_t)(w * 64 + (__ffsll((unsigned long long)stay) - 1)));
				stay &= stay - 1;
			}
		}
		const int na = arrive[k];
		if (na > 0) {
			const int f = first[k];
			for (int r = 0; r < na; ++r) put(dst++, mi[f + r]);
		}
	}
	if (staged) {
		__syncthreads();
		for (long long d = base + threadIdx.x; d < top; d += blockDim.x) perm[d] = stage[d - base];
	}
}

__global__ void k_keys32(Grid g, SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ species_start, uint32_t* __restrict__ keys) {
	long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= P.count) return;
	int i = (int)idx;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	keys[i] = (uint32_t)((((long long)s * g.ncells + P.cell[i]) << 3) | ((P.off[0][i] >= 0.5f) << 2) |
	                     ((P.off[1][i] >= 0.5f) << 1) | (P.off[2][i] >= 0.5f));
}

__global__ void k_permute32(ParticleArrays32 src, ParticleArrays32 dst, const uint32_t* __restrict__ perm) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= src.count) return;
	long long s = perm[i];
	dst.cell[i] = src.cell[s];
#pragma unroll
	for (int c = 0; c < 3; ++c) {
		dst.off[c][i] = src.off[c][s];
		dst.u[c][i] = src.u[c][s];
	}
	dst.id[i] = src.id[s];
}

// FP64 absolute positions and momenta -> cell + offset, u = p / m (owned cells 2..n+1).
__global__ void k_from_fp64(Grid g, SpeciesTable sp, const double* __restrict__ x, const double* __restrict__ p, long long np,
                            const int* __restrict__ species_start, ParticleArrays32 P, int* __restrict__ status) {
	long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= np) return;
	int i = (int)idx;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	int cc[3];
	for (int a = 0; a < 3; ++a) {
		double r = (x[a * np + i] - g.x0[a]) / g.d[a];
		double fl = floor(r);
		int ci = (int)fl;
		float f = (float)(r - fl);
		if (f >= 1.0f) {
			f = 0.0f;
			ci += 1;
		}
		if (ci == g.n[a] + 2) ci = 2;  // x = 2L is the periodic image of x = L
		if (ci < 2 || ci > g.n[a] + 1) atomicOr(status, 4);
		cc[a] = ci < 0 ? 0 : (ci > g.N[a] - 1 ? g.N[a] - 1 : ci);
		P.off[a][i] = f;
		P.u[a][i] = (float)(p[a * np + i] / sp.mass[s]);
	}
	P.cell[i] = (int)cell_id(g, cc[0], cc[1], cc[2]);
}

// cell + offset, u -> FP64 absolute positions and momenta
__global__ void k_to_fp64(Grid g, SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ species_start, double* __restrict__ x,
                          double* __restrict__ p) {
	long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= P.count) return;
	int i = (int)idx;
	long long np = P.count;
	int s = 0;
	while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
	int cc[3];
	cell_ijk(g, P.cell[i], cc);
	for (int a = 0; a < 3; ++a) {
		x[a * np + i] = g.x0[a] + ((double)cc[a] + (double)P.off[a][i]) * g.d[a];
		p[a * np + i] = (double)P.u[a][i] * sp.mass[s];
	}
}

// FP64 component planes [3][n] -> node-major float4 (x, y, z, 0)
__global__ void k_f64_to_f32(const double* __restrict__ a, float4* __restrict__ b, long long n) {
	long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) b[i] = make_float4((float)a[i], (float)a[n + i], (float)a[2 * n + i], 0.0f);
}

// Node / cell values from float corner sums, accumulated in FP64.
__global__ void k_node_assemble32(Grid g, const float* __restrict__ cs, double* __restrict__ J, double* __restrict__ mu) {
	long long nid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (nid >= g.nnodes) return;
	int k = (int)(nid % (g.N[2] + 1));
	int j = (int)((nid / (g.N[2] + 1)) % (g.N[1] + 1));
	int i = (int)(nid / ((long long)(g.N[2] + 1) * (g.N[1] + 1)));
	double s[12];
#pragma unroll
	for (int q = 0; q < 12; ++q) s[q] = 0.0;
	for (int a = 0; a < 2; ++a)
		for (int b = 0; b < 2; ++b)
			for (int cc = 0; cc < 2; ++cc) {
				int ci = i - a, cj = j - b, ck = k - cc;
				if (ci < 0 || cj < 0 || ck < 0 || ci >= g.N[0] || cj >= g.N[1] || ck >= g.N[2]) continue;
				long long cid = cell_id(g, ci, cj, ck);
				int corner = a * 4 + b * 2 + cc;
#pragma unroll
				for (int q = 0; q < 12; ++q) s[q] += (double)cs[((long long)q * 8 + corner) * g.ncells + cid];
			}
#pragma unroll
	for (int q = 0; q < 3; ++q) J[q * g.nnodes + nid] = s[q];
#pragma unroll
	for (int q = 0; q < 9; ++q) mu[q * g.nnodes + nid] = s[3 + q];
}

__global__ void k_cell_assemble32(Grid g, const float* __restrict__ cs, int nq, double* __restrict__ out) {
	long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= g.ncells) return;
	int k = (int)(c % g.N[2]);
	int j = (int)((c / g.N[2]) % g.N[1]);
	int i = (int)(c / ((long long)g.N[2] * g.N[1]));
	for (int q = 0; q < nq; ++q) {
		double s = 0.0;
		for (int a = 0; a < 2; ++a)
			for (int b = 0; b < 2; ++b)
				for (int cc = 0; cc < 2; ++cc) {
					int bi = i - a, bj = j - b, bk = k - cc;
					if (bi < 0 || bj < 0 || bk < 0) continue;
					s += (double)cs[((long long)q * 8 + a * 4 + b * 2 + cc) * g.ncells + cell_id(g, bi, bj, bk)];
				}
		out[(long long)q * g.ncells + c] = s;
	}
}

__global__ void __launch_bounds__(256) k_pressure_reduce32(Grid g, const float* __restrict__ ps, double* __restrict__ Pi) {
	int i = blockIdx.x / 4;
	int tj = (blockIdx.x / 2) % 2;
	int tk = blockIdx.x % 2;
	__shared__ double red[6][256];
	double acc[6] = {0, 0, 0, 0, 0, 0};
	long long plane = (long long)g.N[1] * g.N[2];
	for (int ti = 0; ti < 2; ++ti) {
		int bi = i - ti;
		if (bi < 0 || bi >= g.N[0]) continue;
		int t = ti * 4 + tj * 2 + tk;
		for (long long r = threadIdx.x; r < plane; r += blockDim.x) {
			long long b = (long long)bi * plane + r;
#pragma unroll
			for (int q = 0; q < 6; ++q) acc[q] += (double)ps[((long long)t * 6 + q) * g.ncells + b];
		}
	}
#pragma unroll
	for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
#pragma unroll
			for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		long long c = cell_id(g, i, tj, tk);
		const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
		for (int q = 0; q < 9; ++q) Pi[q * g.ncells + c] = red[sym[q]][0];
	}
}

// kinetic energy with the compensated gamma - 1 = u^2 / (gamma + 1), and momentum
__global__ void __launch_bounds__(256) k_particle_energy32(SpeciesTable sp, ParticleArrays32 P, const int* __restrict__ species_start,
                                                           double* __restrict__ partial) {
	__shared__ double red[4][256];
	double acc[4] = {0, 0, 0, 0};
	for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.count; i += (long long)gridDim.x * blockDim.x) {
		int s = 0;
		while (s + 1 < sp.count && i >= species_start[s + 1]) ++s;
		double m = sp.mass[s], w = sp.weight[s];
		float u[3] = {P.u[0][i], P.u[1][i], P.u[2][i]};
		float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
		float gm1 = u2 / (sqrtf(1.0f + u2) + 1.0f);
		acc[0] += (double)gm1 * m * w;
		for (int c = 0; c < 3; ++c) acc[1 + c] += (double)u[c] * m * w;
	}
	for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = acc[q];
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
			for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int q = 0; q < 4; ++q) partial[q * gridDim.x + blockIdx.x] = red[q][0];
}

// ------------------------------------------------------------------ launch wrappers
namespace {
inline unsigned blocks32(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

// Lanes per cell group of the moment deposits at x^n (node and cell moments). Measured at
// 128^3 with 2 species x 64 ppc: 8 lanes (16-cell tiles, 128-lane workgroups) beat 16 and 32.
constexpr int kMomentLanes = 8;

int group_lanes(int nspecies) {
	int G = 8;
	while (G < 8 * nspecies) G *= 2;
	return G;
}

// Deposit workgroups: T cells of one z row (as many as fit 256 lanes, at most kMaxTile).
int deposit_tile(int G) { return 256 / G < kMaxTile ? 256 / G : kMaxTile; }
unsigned deposit_blocks(const Grid& g, int T) { return (unsigned)(g.N[0] * g.N[1] * ((g.N[2] + T - 1) / T)); }
size_t box_bytes(int T) { return (size_t)9 * (T + 2) * sizeof(float4); }

void launch_node_moments32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, const float* E,
                           const float* B, float dt, float theta, float* scratch, hipStream_t s) {
	int G = kMomentLanes, T = deposit_tile(G);
	hipLaunchKernelGGL(k_node_moments32, dim3(deposit_blocks(g, T)), dim3(T * G), box_bytes(T), s, g, sp, P, start,
	                   reinterpret_cast<const float4*>(E), reinterpret_cast<const float4*>(B), dt, theta, G, T, scratch);
}
// Dual-cell plane tiles: Tk = 4 bases along z (fewer when T < 4), Tj = T / Tk along y; the
// grid is a multiple of 8 blocks (one run of tiles per XCD).
struct PlaneTiling {
	int Tj, Tk;
	unsigned blocks;
};
PlaneTiling plane_tiling(const Grid& g, int T) {
	PlaneTiling p;
	p.Tk = T >= 4 ? 4 : T;
	p.Tj = T / p.Tk;
	long long n = (long long)g.N[0] * ((g.N[1] + p.Tj - 1) / p.Tj) * ((g.N[2] + p.Tk - 1) / p.Tk);
	p.blocks = (unsigned)(8 * ((n + 7) / 8));
	return p;
}

void launch_cell_moments32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, const float* E,
                           const float* B, float dt, float* rho_scratch, float* pi_scratch, hipStream_t s, float* dens_scratch) {
	int G = kMomentLanes, T = deposit_tile(G);
	PlaneTiling pt = plane_tiling(g, T);
	size_t box = (size_t)3 * (pt.Tj + 2) * (pt.Tk + 2) * sizeof(float4);
	hipLaunchKernelGGL(dens_scratch ? k_cell_moments32<true> : k_cell_moments32<false>, dim3(pt.blocks), dim3(T * G), box, s, g, sp, P,
	                   start, reinterpret_cast<const float4*>(E), reinterpret_cast<const float4*>(B), dt, G, pt.Tj, pt.Tk, rho_scratch,
	                   pi_scratch, dens_scratch);
}
void launch_density32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* start, float* scratch,
                      hipStream_t s) {
	// Lanes per dual cell: with many particles per cell the wider groups win (128 particles:
	// 16 lanes 3.96 ms against 8 lanes 4.26 ms at 128^3), with few the per-cell work dominates
	// (32 particles: 8 lanes 3.20 ms against 16 lanes 4.69 ms in the 1024 x 64 x 64 shock).
	int G = (double)P.count / (double)g.ncells >= 64.0 ? group_lanes(sp.count) : 8;
	int T = deposit_tile(G);
	PlaneTiling pt = plane_tiling(g, T);
	hipLaunchKernelGGL(k_density32, dim3(pt.blocks), dim3(T * G), 0, s, g, sp, P, start, G, pt.Tj, pt.Tk, scratch);
}
static void magic_divisor(uint32_t d, uint32_t& m, int& l) {
	l = 0;
	while ((1ull << l) < d) ++l;
	m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
}

void launch_push32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, const float* E,
                   const float* Et, const float* B, float dt, float eta, uint32_t* keys, uint64_t* mask, int* status, int* dead,
                   uint32_t* esc, uint64_t* bounce_mask, hipStream_t s) {
	if (P.count == 0) return;
	Push32Consts K{};
	for (int q = 0; q < sp.count; ++q) {  // the same float arithmetic as q dt / m on the device
		const float mf = (float)sp.mass[q], qf = (float)sp.charge[q];
		K.qdm[q] = qf * dt / mf;
		K.beta[q] = 0.5f * qf * dt / mf;
	}
	for (int a = 0; a < 3; ++a) {
		K.inv_d[a] = (float)(1.0 / g.d[a]);
		K.d[a] = (float)g.d[a];
	}
	magic_divisor((uint32_t)g.N[2], K.m2, K.l2);  // N >= 5: l >= 3
	magic_divisor((uint32_t)g.N[1], K.m1, K.l1);
	hipLaunchKernelGGL(g.wall ? k_push32<true> : k_push32<false>, dim3(blocks32(P.count, 256)), dim3(256), 0, s, g, sp, K, P, species_start,
	                   reinterpret_cast<const float4*>(E), reinterpret_cast<const float4*>(Et), reinterpret_cast<const float4*>(B), dt,
	                   eta, keys, mask, status, dead, esc, bounce_mask);
}
void launch_mask_counts(const uint64_t* mask, long long nwords, int* counts, hipStream_t s) {
	hipLaunchKernelGGL(k_mask_counts, dim3(blocks32(nwords + 1, 256)), dim3(256), 0, s, mask, nwords, counts);
}
void launch_mark_movers(uint64_t* mask, long long lo, long long hi, hipStream_t s) {
	if (hi <= lo) return;
	const long long words = ((hi - 1) >> 6) - (lo >> 6) + 1;
	hipLaunchKernelGGL(k_mark_movers, dim3(blocks32(words, 256)), dim3(256), 0, s, mask, lo, hi);
}
void launch_gather_movers(const uint64_t* mask, const int* prefix, const uint32_t* keys, long long n, long long nkeys, uint32_t* mk,
                          uint32_t* mi, hipStream_t s) {
	if (n == 0) return;
	const long long nwords = (n + 63) / 64;
	hipLaunchKernelGGL(k_gather_movers, dim3(blocks32(nwords, 256)), dim3(256), 0, s, mask, prefix, keys, n, nwords, (uint32_t)nkeys,
	                   mk, mi);
}
void launch_first_of_run(const uint32_t* mk, long long M, int* first, hipStream_t s) {
	if (M == 0) return;
	hipLaunchKernelGGL(k_first_of_run, dim3(blocks32(M, 256)), dim3(256), 0, s, mk, M, first);
}
void launch_rebin_counts(const int* start_old, const uint64_t* mask, const int* prefix, const int* arrive, long long nkeys, int* counts,
                         hipStream_t s) {
	hipLaunchKernelGGL(k_rebin_counts, dim3(blocks32(nkeys + 1, 256)), dim3(256), 0, s, start_old, mask, prefix, arrive, nkeys, counts);
}
void launch_rebin_place(const int* start_old, const int* start_new, const uint64_t* mask, const int* arrive, const int* first,
                        const uint32_t* mi, long long nkeys, uint32_t* perm, hipStream_t s) {
	hipLaunchKernelGGL(k_rebin_place, dim3(blocks32(nkeys, 256)), dim3(256), 0, s, start_old, start_new, mask, arrive, first, mi, nkeys,
	                   perm);
}
void launch_keys32(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, uint32_t* keys,
                   hipStream_t s) {
	if (P.count == 0) return;
	hipLaunchKernelGGL(k_keys32, dim3(blocks32(P.count, 256)), dim3(256), 0, s, g, sp, P, species_start, keys);
}
void launch_permute32(const ParticleArrays32& src, const ParticleArrays32& dst, const uint32_t* perm, hipStream_t s) {
	if (src.count == 0) return;
	hipLaunchKernelGGL(k_permute32, dim3(blocks32(src.count, 256)), dim3(256), 0, s, src, dst, perm);
}
void launch_from_fp64(const Grid& g, const SpeciesTable& sp, const double* x, const double* p, long long np,
                      const int* species_start, const ParticleArrays32& P, int* status, hipStream_t s) {
	if (np == 0) return;
	hipLaunchKernelGGL(k_from_fp64, dim3(blocks32(np, 256)), dim3(256), 0, s, g, sp, x, p, np, species_start, P, status);
}
void launch_to_fp64(const Grid& g, const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, double* x,
                    double* p, hipStream_t s) {
	if (P.count == 0) return;
	hipLaunchKernelGGL(k_to_fp64, dim3(blocks32(P.count, 256)), dim3(256), 0, s, g, sp, P, species_start, x, p);
}
void launch_f64_to_f32(const double* a, float* b, long long n, hipStream_t s) {
	hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks32(n, 256)), dim3(256), 0, s, a, reinterpret_cast<float4*>(b), n);
}
void launch_node_assemble32(const Grid& g, const float* scratch, double* J, double* mu, hipStream_t s) {
	hipLaunchKernelGGL(k_node_assemble32, dim3(blocks32(g.nnodes, 256)), dim3(256), 0, s, g, scratch, J, mu);
}
void launch_cell_assemble32(const Grid& g, const float* scratch, int nq, double* out, hipStream_t s) {
	hipLaunchKernelGGL(k_cell_assemble32, dim3(blocks32(g.ncells, 256)), dim3(256), 0, s, g, scratch, nq, out);
}
void launch_pressure_reduce32(const Grid& g, const float* pi_scratch, double* Pi, hipStream_t s) {
	hipLaunchKernelGGL(k_pressure_reduce32, dim3(g.N[0] * 4), dim3(256), 0, s, g, pi_scratch, Pi);
}
void launch_particle_energy32(const SpeciesTable& sp, const ParticleArrays32& P, const int* species_start, double* partial,
                              hipStream_t s) {
	hipLaunchKernelGGL(k_particle_energy32, dim3(kEnergyBlocks), dim3(256), 0, s, sp, P, species_start, partial);
}

}  // namespace picgpu
