/*
 * Learned block preconditioner of the transport solve. A per-item network
 * maps the 7 numbers that fix the diagonal block
 *   D = I + tau Minv F_ee,  F_ee = -sum_k a_k INT_k + sum_{f outflow} s_f L_f S L_f^T
 * (x = tau a_0..2, tau s_0..3 with s_f = 0 on non-outflow faces, tau =
 * theta dt / det, a_k = v . gamma_k) to a correction of the identity,
 * P = I + NN(x) ~ D^-1. The block solve D u = r then runs as preconditioned
 * Richardson u <- u + P (r - D u) to a relative residual of 1e-8, with the
 * exact inverse (the stored Gauss-Jordan factor) as the fallback and as the
 * yardstick of every result.
 *
 * Network: 7 -> 64 -> 64 -> 100, tanh, inputs and outputs normalised by
 * training-set means and deviations. Products on MFMA in bf16 with fp32
 * accumulation (v_mfma_f32_16x16x16_bf16): one wave = 16 items; lane l holds
 * A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and D[4(l/16)+j][l%16], j < 4. A
 * result tile in D layout is therefore directly the A operand of its
 * transpose and the B operand along the item dimension, which is all the
 * weight gradients need; the activations go through LDS where they are the
 * A operand of the next layer. Training: per-wave partial gradients, a
 * fixed-order sum and Adam in one kernel, fp32 master weights.
 */
#include "ctx.h"
#include "learned_block.h"

using namespace gdh;

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define MLP_IN 7
#define MLP_H 64
#define MLP_OUT 100
#define MLP_OUTP 112 /* padded to 7 tiles */
/* flat fp32 parameters: W1[64][16] b1[64] W2[64][64] b2[64] W3[112][64] b3[112] */
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + 64 * 16)
#define OFF_W2 (OFF_B1 + 64)
#define OFF_B2 (OFF_W2 + 64 * 64)
#define OFF_W3 (OFF_B2 + 64)
#define OFF_B3 (OFF_W3 + MLP_OUTP * 64)
#define NPARAM (OFF_B3 + MLP_OUTP)

struct gdh_mlp {
	float *w;	 /* fp32 master parameters (device) */
	short *wb;	 /* bf16 copies of W1, W2, W3 (same offsets) */
	float *m, *v;	 /* Adam moments */
	float *gpart;	 /* per-wave partial gradients */
	int gpart_waves;
	float *norm;	 /* mu_x[7] sd_x[7] mu_y[100] sd_y[100] */
	double *loss;	 /* per-wave partial losses */
	int step;
};

__device__ __forceinline__ short to_bf16(float x)
{
	unsigned u = __float_as_uint(x);
	u += 0x7fff + ((u >> 16) & 1); /* round to nearest even */
	return (short)(u >> 16);
}

__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

/* B operand of W^T for a weight matrix W[nout][nin] (bf16, row-major):
 * B[k][n] = W[n0 + n][k0 + k] */
__device__ __forceinline__ bf4 b_wt(const short *W, int nin, int k0, int n0)
{
	const int l = threadIdx.x & 63;
	bf4 b;
	for (int j = 0; j < 4; j++)
		b[j] = W[(n0 + l % 16) * nin + k0 + 4 * (l / 16) + j];
	return b;
}

/* B operand of W for W[nout][nin]: B[k][n] = W[k0 + k][n0 + n] */
__device__ __forceinline__ bf4 b_w(const short *W, int nin, int k0, int n0)
{
	const int l = threadIdx.x & 63;
	bf4 b;
	for (int j = 0; j < 4; j++)
		b[j] = W[(k0 + 4 * (l / 16) + j) * nin + n0 + l % 16];
	return b;
}

/* a D-layout tile (16 items x 16) as a bf16 operand (A of its transpose, or
 * B along the item dimension) */
__device__ __forceinline__ bf4 as_bf(f4 d)
{
	bf4 r;
	for (int j = 0; j < 4; j++)
		r[j] = to_bf16(d[j]);
	return r;
}

/* store a D-layout tile (items x cols) into the wave's LDS matrix S[16][ld] */
__device__ __forceinline__ void st_d(float *S, int ld, int c0, f4 d)
{
	const int l = threadIdx.x & 63;
	for (int j = 0; j < 4; j++)
		S[(4 * (l / 16) + j) * ld + c0 + l % 16] = d[j];
}

/* A operand from the wave's LDS matrix S[16 items][ld], columns k0.. */
__device__ __forceinline__ bf4 a_lds(const float *S, int ld, int k0)
{
	const int l = threadIdx.x & 63;
	bf4 a;
	for (int j = 0; j < 4; j++)
		a[j] = to_bf16(S[(l % 16) * ld + k0 + 4 * (l / 16) + j]);
	return a;
}

struct Fwd {
	f4 h1[4], h2[4], y[7]; /* D layout, after tanh for h1 h2; y normalised */
};

/* forward pass of the wave's 16 items; x normalised, rows i < n valid */
__device__ inline void mlp_forward(const float *__restrict__ w, const short *__restrict__ wb,
				   const float (&xn)[4], float *S, Fwd &F)
{
	const int l = threadIdx.x & 63;
	bf4 ax;
	for (int j = 0; j < 4; j++)
		ax[j] = to_bf16(xn[j]);
	for (int t = 0; t < 4; t++) {
		f4 c = { 0, 0, 0, 0 };
		c = mfma(ax, b_wt(wb + OFF_W1, 16, 0, 16 * t), c);
		for (int j = 0; j < 4; j++)
			c[j] = tanhf(c[j] + w[OFF_B1 + 16 * t + l % 16]);
		F.h1[t] = c;
		st_d(S, 64, 16 * t, c);
	}
	__syncthreads();
	bf4 a1[4];
	for (int kc = 0; kc < 4; kc++)
		a1[kc] = a_lds(S, 64, 16 * kc);
	__syncthreads();
	for (int t = 0; t < 4; t++) {
		f4 c = { 0, 0, 0, 0 };
		for (int kc = 0; kc < 4; kc++)
			c = mfma(a1[kc], b_wt(wb + OFF_W2, 64, 16 * kc, 16 * t), c);
		for (int j = 0; j < 4; j++)
			c[j] = tanhf(c[j] + w[OFF_B2 + 16 * t + l % 16]);
		F.h2[t] = c;
		st_d(S, 64, 16 * t, c);
	}
	__syncthreads();
	bf4 a2[4];
	for (int kc = 0; kc < 4; kc++)
		a2[kc] = a_lds(S, 64, 16 * kc);
	__syncthreads();
	for (int t = 0; t < 7; t++) {
		f4 c = { 0, 0, 0, 0 };
		for (int kc = 0; kc < 4; kc++)
			c = mfma(a2[kc], b_wt(wb + OFF_W3, 64, 16 * kc, 16 * t), c);
		for (int j = 0; j < 4; j++)
			c[j] += w[OFF_B3 + 16 * t + l % 16];
		F.y[t] = c;
	}
}

/* this lane's 4 normalised inputs of the A layout (item l%16, features
 * 4(l/16)+j) and of the D layout (item 4(l/16)+j, feature l%16) */
__device__ inline void load_x(const float *__restrict__ x, const float *__restrict__ norm,
			      int i0, int n, float (&xa)[4], float (&xd)[4])
{
	const int l = threadIdx.x & 63;
	for (int j = 0; j < 4; j++) {
		const int ia = i0 + l % 16, fa = 4 * (l / 16) + j;
		xa[j] = (ia < n && fa < MLP_IN) ? (x[(long)ia * MLP_IN + fa] - norm[fa]) / norm[7 + fa] : 0.f;
		const int id = i0 + 4 * (l / 16) + j, fd = l % 16;
		xd[j] = (id < n && fd < MLP_IN) ? (x[(long)id * MLP_IN + fd] - norm[fd]) / norm[7 + fd] : 0.f;
	}
}

/* One wave per 16 items: loss and this wave's gradient of
 * mean over items and outputs of (y_pred - y_norm)^2 (scaled by 1/ntot). */
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ w,
						  const short *__restrict__ wb,
						  const float *__restrict__ norm,
						  const float *__restrict__ x,
						  const float *__restrict__ y, int n, long ntot,
						  float *__restrict__ gpart,
						  double *__restrict__ loss)
{
	__shared__ float S_all[4][16 * MLP_OUTP];
	float *S = S_all[threadIdx.x / 64];
	const int l = threadIdx.x & 63;
	const int wave = blockIdx.x * 4 + threadIdx.x / 64;
	const int i0 = wave * 16;
	float xa[4], xd[4];
	load_x(x, norm, i0, n, xa, xd);
	Fwd F;
	mlp_forward(w, wb, xa, S, F);
	/* dY (D layout), loss */
	f4 dy[7];
	double ls = 0.0;
	const float scale = 2.0f / (float)(ntot * MLP_OUT);
	for (int t = 0; t < 7; t++)
		for (int j = 0; j < 4; j++) {
			const int it = i0 + 4 * (l / 16) + j, o = 16 * t + l % 16;
			float d = 0.f;
			if (it < n && o < MLP_OUT) {
				const float yt = (y[(long)it * MLP_OUT + o] - norm[14 + o]) / norm[114 + o];
				d = F.y[t][j] - yt;
				ls += (double)d * d;
			}
			dy[t][j] = d * scale;
		}
	for (int s = 32; s > 0; s >>= 1)
		ls += __shfl_xor(ls, s);
	float *g = gpart + (long)wave * NPARAM;
	/* layer 3: dW3[o][k] = sum_items dY[i][o] h2[i][k]; db3 */
	for (int t = 0; t < 7; t++) {
		const bf4 at = as_bf(dy[t]);
		for (int kc = 0; kc < 4; kc++) {
			f4 c = { 0, 0, 0, 0 };
			c = mfma(at, as_bf(F.h2[kc]), c);
			for (int j = 0; j < 4; j++)
				g[OFF_W3 + (16 * t + 4 * (l / 16) + j) * 64 + 16 * kc + l % 16] = c[j];
		}
		float sb = dy[t][0] + dy[t][1] + dy[t][2] + dy[t][3];
		sb += __shfl_xor(sb, 16);
		sb += __shfl_xor(sb, 32);
		if (l < 16)
			g[OFF_B3 + 16 * t + l] = sb;
	}
	/* dH2 = dY W3 (A = dY in A layout via LDS) */
	for (int t = 0; t < 7; t++)
		st_d(S, MLP_OUTP, 16 * t, dy[t]);
	__syncthreads();
	bf4 ady[7];
	for (int kc = 0; kc < 7; kc++)
		ady[kc] = a_lds(S, MLP_OUTP, 16 * kc);
	__syncthreads();
	f4 dz2[4];
	for (int t = 0; t < 4; t++) {
		f4 c = { 0, 0, 0, 0 };
		for (int kc = 0; kc < 7; kc++)
			c = mfma(ady[kc], b_w(wb + OFF_W3, 64, 16 * kc, 16 * t), c);
		for (int j = 0; j < 4; j++)
			c[j] *= 1.f - F.h2[t][j] * F.h2[t][j];
		dz2[t] = c;
	}
	/* layer 2: dW2[o][k] = sum dZ2[i][o] h1[i][k]; db2 */
	for (int t = 0; t < 4; t++) {
		const bf4 at = as_bf(dz2[t]);
		for (int kc = 0; kc < 4; kc++) {
			f4 c = { 0, 0, 0, 0 };
			c = mfma(at, as_bf(F.h1[kc]), c);
			for (int j = 0; j < 4; j++)
				g[OFF_W2 + (16 * t + 4 * (l / 16) + j) * 64 + 16 * kc + l % 16] = c[j];
		}
		float sb = dz2[t][0] + dz2[t][1] + dz2[t][2] + dz2[t][3];
		sb += __shfl_xor(sb, 16);
		sb += __shfl_xor(sb, 32);
		if (l < 16)
			g[OFF_B2 + 16 * t + l] = sb;
	}
	/* dH1 = dZ2 W2 */
	for (int t = 0; t < 4; t++)
		st_d(S, 64, 16 * t, dz2[t]);
	__syncthreads();
	bf4 adz[4];
	for (int kc = 0; kc < 4; kc++)
		adz[kc] = a_lds(S, 64, 16 * kc);
	__syncthreads();
	bf4 bx;
	for (int j = 0; j < 4; j++)
		bx[j] = to_bf16(xd[j]);
	for (int t = 0; t < 4; t++) {
		f4 c = { 0, 0, 0, 0 };
		for (int kc = 0; kc < 4; kc++)
			c = mfma(adz[kc], b_w(wb + OFF_W2, 64, 16 * kc, 16 * t), c);
		for (int j = 0; j < 4; j++)
			c[j] *= 1.f - F.h1[t][j] * F.h1[t][j];
		/* layer 1: dW1[o][f] = sum dZ1[i][o] x[i][f]; db1 */
		f4 d = { 0, 0, 0, 0 };
		d = mfma(as_bf(c), bx, d);
		for (int j = 0; j < 4; j++)
			g[OFF_W1 + (16 * t + 4 * (l / 16) + j) * 16 + l % 16] = d[j];
		float sb = c[0] + c[1] + c[2] + c[3];
		sb += __shfl_xor(sb, 16);
		sb += __shfl_xor(sb, 32);
		if (l < 16)
			g[OFF_B1 + 16 * t + l] = sb;
	}
	if (l == 0)
		loss[wave] = ls;
}

/* fixed-order sum of the wave partials; Adam step when lr > 0, else the
 * summed gradient goes to gout */
__global__ void k_mlp_update(float *__restrict__ w, short *__restrict__ wb,
			     float *__restrict__ m, float *__restrict__ v,
			     const float *__restrict__ gpart, int nw, float lr, int step,
			     float *__restrict__ gout)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NPARAM)
		return;
	float gs = 0.f;
	for (int k = 0; k < nw; k++)
		gs += gpart[(long)k * NPARAM + p];
	/* padded rows of W3 / b3 and padded input columns of W1 stay zero */
	const bool pad = (p >= OFF_W3 + MLP_OUT * 64 && p < OFF_B3) || p >= OFF_B3 + MLP_OUT ||
			 (p < OFF_B1 && (p % 16) >= MLP_IN);
	if (pad)
		gs = 0.f;
	if (gout) {
		gout[p] = gs;
		return;
	}
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float mm = b1 * m[p] + (1 - b1) * gs, vv = b2 * v[p] + (1 - b2) * gs * gs;
	m[p] = mm;
	v[p] = vv;
	const float mh = mm / (1 - powf(b1, (float)step)), vh = vv / (1 - powf(b2, (float)step));
	const float nw_ = pad ? 0.f : w[p] - lr * mh / (sqrtf(vh) + eps);
	w[p] = nw_;
	wb[p] = to_bf16(nw_);
}

/* predictions of n items, denormalised: y[i][o] (D^-1 - I) */
__global__ __launch_bounds__(256) void k_mlp_predict(const float *__restrict__ w,
						     const short *__restrict__ wb,
						     const float *__restrict__ norm,
						     const float *__restrict__ x, int n,
						     float *__restrict__ y)
{
	__shared__ float S_all[4][16 * 64];
	float *S = S_all[threadIdx.x / 64];
	const int l = threadIdx.x & 63;
	const int i0 = (blockIdx.x * 4 + threadIdx.x / 64) * 16;
	float xa[4], xd[4];
	load_x(x, norm, i0, n, xa, xd);
	Fwd F;
	mlp_forward(w, wb, xa, S, F);
	for (int t = 0; t < 7; t++)
		for (int j = 0; j < 4; j++) {
			const int it = i0 + 4 * (l / 16) + j, o = 16 * t + l % 16;
			if (it < n && o < MLP_OUT)
				y[(long)it * MLP_OUT + o] = F.y[t][j] * norm[114 + o] + norm[14 + o];
		}
}

__global__ void k_gather_rows(const float *__restrict__ x, const float *__restrict__ y,
			      const int *__restrict__ idx, int batch,
			      float *__restrict__ bx, float *__restrict__ by)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= batch * (MLP_IN + MLP_OUT))
		return;
	const int b = k / (MLP_IN + MLP_OUT), c = k % (MLP_IN + MLP_OUT);
	if (c < MLP_IN)
		bx[(long)b * MLP_IN + c] = x[(long)idx[b] * MLP_IN + c];
	else
		by[(long)b * MLP_OUT + c - MLP_IN] = y[(long)idx[b] * MLP_OUT + c - MLP_IN];
}

/* ------------------------------------------------------------------ */

static void mlp_sync_bf16(gdh_mlp *p);

__global__ void k_to_bf16(const float *__restrict__ w, short *__restrict__ wb, int n)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n)
		wb[k] = to_bf16(w[k]);
}

static void mlp_sync_bf16(gdh_mlp *p)
{
	k_to_bf16<<<nblocks(NPARAM, 256), 256>>>(p->w, p->wb, NPARAM);
	HIPCHECK(hipGetLastError());
}

extern "C" int gdh_mlp_nparams(void)
{
	return NPARAM;
}

/* Xavier-uniform weights from SplitMix64(seed), zero biases and padding */
extern "C" gdh_mlp *gdh_mlp_create(uint64_t seed)
{
	gdh_mlp *p = new gdh_mlp();
	std::vector<float> h(NPARAM, 0.f);
	uint64_t s = seed;
	auto uni = [&s]() {
		uint64_t z = (s += 0x9e3779b97f4a7c15ull);
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
		z ^= z >> 31;
		return (double)(z >> 11) * (1.0 / 9007199254740992.0);
	};
	auto fill = [&](int off, int rows, int cols, int nin, int nout) {
		const double a = sqrt(6.0 / (nin + nout));
		for (int r = 0; r < rows; r++)
			for (int c = 0; c < cols; c++)
				h[off + r * cols + c] = (float)((2 * uni() - 1) * a);
	};
	fill(OFF_W1, 64, MLP_IN, MLP_IN, 64); /* rows of 16, first 7 used */
	for (int r = 63; r >= 0; r--) /* spread the 7 columns to the stride 16 */
		for (int c = 15; c >= 0; c--)
			h[OFF_W1 + r * 16 + c] = c < MLP_IN ? h[OFF_W1 + r * MLP_IN + c] : 0.f;
	fill(OFF_W2, 64, 64, 64, 64);
	fill(OFF_W3, MLP_OUT, 64, 64, MLP_OUT);
	p->w = (float *)dmalloc(sizeof(float) * NPARAM);
	p->wb = (short *)dmalloc(sizeof(short) * NPARAM);
	p->m = (float *)dmalloc(sizeof(float) * NPARAM);
	p->v = (float *)dmalloc(sizeof(float) * NPARAM);
	p->norm = (float *)dmalloc(sizeof(float) * 214);
	HIPCHECK(hipMemcpy(p->w, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
	HIPCHECK(hipMemset(p->m, 0, sizeof(float) * NPARAM));
	HIPCHECK(hipMemset(p->v, 0, sizeof(float) * NPARAM));
	std::vector<float> nrm(214, 0.f);
	for (int k = 0; k < 7; k++)
		nrm[7 + k] = 1.f;
	for (int k = 0; k < 100; k++)
		nrm[114 + k] = 1.f;
	HIPCHECK(hipMemcpy(p->norm, nrm.data(), sizeof(float) * 214, hipMemcpyHostToDevice));
	p->gpart = nullptr;
	p->gpart_waves = 0;
	p->loss = nullptr;
	p->step = 0;
	mlp_sync_bf16(p);
	HIPCHECK(hipDeviceSynchronize());
	return p;
}

extern "C" void gdh_mlp_destroy(gdh_mlp *p)
{
	if (!p)
		return;
	HIPCHECK(hipDeviceSynchronize());
	for (void *q : { (void *)p->w, (void *)p->wb, (void *)p->m, (void *)p->v,
			 (void *)p->gpart, (void *)p->norm, (void *)p->loss })
		if (q)
			HIPCHECK(hipFree(q));
	delete p;
}

extern "C" void gdh_mlp_get_params(gdh_mlp *p, float *out)
{
	HIPCHECK(hipMemcpy(out, p->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
}

extern "C" void gdh_mlp_set_params(gdh_mlp *p, const float *in)
{
	HIPCHECK(hipMemcpy(p->w, in, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
	mlp_sync_bf16(p);
	HIPCHECK(hipDeviceSynchronize());
}

/* mu_x[7], sd_x[7], mu_y[100], sd_y[100] */
extern "C" void gdh_mlp_set_normalization(gdh_mlp *p, const float *norm214)
{
	HIPCHECK(hipMemcpy(p->norm, norm214, sizeof(float) * 214, hipMemcpyHostToDevice));
}

static void ensure_parts(gdh_mlp *p, int waves)
{
	if (p->gpart_waves >= waves)
		return;
	if (p->gpart)
		HIPCHECK(hipFree(p->gpart));
	if (p->loss)
		HIPCHECK(hipFree(p->loss));
	p->gpart = (float *)dmalloc(sizeof(float) * NPARAM * (size_t)waves);
	p->loss = (double *)dmalloc(sizeof(double) * waves);
	p->gpart_waves = waves;
}

/* gradient and loss of the batch (device arrays x[n][7], y[n][100]);
 * with lr > 0 an Adam step is taken, else the gradient goes to gout */
static double mlp_step(gdh_mlp *p, const float *dx, const float *dy, int n, float lr,
		       float *gout_dev)
{
	const int waves = (n + 15) / 16, blocks = (waves + 3) / 4;
	ensure_parts(p, blocks * 4);
	HIPCHECK(hipMemset(p->gpart, 0, sizeof(float) * NPARAM * (size_t)blocks * 4));
	HIPCHECK(hipMemset(p->loss, 0, sizeof(double) * blocks * 4));
	k_mlp_grad<<<blocks, 256>>>(p->w, p->wb, p->norm, dx, dy, n, n, p->gpart, p->loss);
	HIPCHECK(hipGetLastError());
	if (lr > 0)
		p->step++;
	k_mlp_update<<<nblocks(NPARAM, 256), 256>>>(p->w, p->wb, p->m, p->v, p->gpart,
						    blocks * 4, lr, p->step, gout_dev);
	HIPCHECK(hipGetLastError());
	std::vector<double> ls(blocks * 4);
	HIPCHECK(hipMemcpy(ls.data(), p->loss, sizeof(double) * ls.size(), hipMemcpyDeviceToHost));
	double s = 0;
	for (double v : ls)
		s += v;
	return s / ((double)n * MLP_OUT);
}

/* Full-batch gradient of the mean squared normalised error (host arrays);
 * returns the loss. */
extern "C" double gdh_mlp_grad(gdh_mlp *p, const float *x, const float *y, int n,
			       float *grad)
{
	float *dx = (float *)dmalloc(sizeof(float) * MLP_IN * (size_t)n);
	float *dy = (float *)dmalloc(sizeof(float) * MLP_OUT * (size_t)n);
	float *dg = (float *)dmalloc(sizeof(float) * NPARAM);
	HIPCHECK(hipMemcpy(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(dy, y, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice));
	const double loss = mlp_step(p, dx, dy, n, 0.f, dg);
	HIPCHECK(hipMemcpy(grad, dg, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
	for (void *q : { (void *)dx, (void *)dy, (void *)dg })
		HIPCHECK(hipFree(q));
	return loss;
}

/* Adam on minibatches of `batch` items drawn (SplitMix64(seed)) from the n
 * training items; loss_hist[k] = loss of minibatch k (may be NULL).
 * Returns the last minibatch loss. */
extern "C" double gdh_mlp_train(gdh_mlp *p, const float *x, const float *y, int n,
				int steps, int batch, float lr, uint64_t seed,
				double *loss_hist)
{
	float *dx = (float *)dmalloc(sizeof(float) * MLP_IN * (size_t)n);
	float *dy = (float *)dmalloc(sizeof(float) * MLP_OUT * (size_t)n);
	float *bx = (float *)dmalloc(sizeof(float) * MLP_IN * (size_t)batch);
	float *by = (float *)dmalloc(sizeof(float) * MLP_OUT * (size_t)batch);
	int *didx = (int *)dmalloc(sizeof(int) * (size_t)batch);
	HIPCHECK(hipMemcpy(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(dy, y, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice));
	std::vector<int> pick(batch);
	uint64_t s = seed;
	double last = 0;
	for (int k = 0; k < steps; k++) {
		for (int b = 0; b < batch; b++) {
			uint64_t z = (s += 0x9e3779b97f4a7c15ull);
			z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
			z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
			z ^= z >> 31;
			pick[b] = (int)(z % (uint64_t)n);
		}
		HIPCHECK(hipMemcpy(didx, pick.data(), sizeof(int) * batch, hipMemcpyHostToDevice));
		k_gather_rows<<<nblocks((long)batch * (MLP_IN + MLP_OUT), 256), 256>>>(
			dx, dy, didx, batch, bx, by);
		HIPCHECK(hipGetLastError());
		last = mlp_step(p, bx, by, batch, lr, nullptr);
		if (loss_hist)
			loss_hist[k] = last;
	}
	for (void *q : { (void *)dx, (void *)dy, (void *)bx, (void *)by, (void *)didx })
		HIPCHECK(hipFree(q));
	return last;
}

extern "C" void gdh_mlp_predict(gdh_mlp *p, const float *x, int n, float *y)
{
	float *dx = (float *)dmalloc(sizeof(float) * MLP_IN * (size_t)n);
	float *dy = (float *)dmalloc(sizeof(float) * MLP_OUT * (size_t)n);
	HIPCHECK(hipMemcpy(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
	const int blocks = ((n + 15) / 16 + 3) / 4;
	k_mlp_predict<<<blocks, 256>>>(p->w, p->wb, p->norm, dx, n, dy);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpy(y, dy, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost));
	HIPCHECK(hipFree(dx));
	HIPCHECK(hipFree(dy));
}

/* ------------------------------------------------------------------ */
/* The network in the transport block solves of a kinetic context       */
/* ------------------------------------------------------------------ */

#define THETA_L 0.5 /* Crank-Nicolson (src/dg/relaxation.c:133) */

/* outward flux coefficient v . vn(e, f) */
__device__ inline double flux_coef(const double ct[9], int f, const double *v)
{
	double vn[3];
	face_normal(ct, f, vn);
	return v[0] * vn[0] + v[1] * vn[1] + v[2] * vn[2];
}

/* D = I + theta dt/det Minv F_ee(v) of item (v, e) and its 7 features, with
 * the formulas of the block factorisation (src/simulation/simulation.c:
 * 298-308, 531-562) */
__device__ inline void item_block(const double *__restrict__ geom,
				  const uint8_t *__restrict__ cls, const KinPar &kp,
				  double dt, int v, int e, double D[10][10], float feat[7])
{
	double ct[9];
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double det = geom[10L * e + 9];
	const double *vv = kp.vi + 3 * v;
	const double tau = THETA_L * dt / det;
	double a[3], sp[4];
	for (int k = 0; k < 3; k++)
		a[k] = ct[0 + k] * vv[0] + ct[3 + k] * vv[1] + ct[6 + k] * vv[2];
	for (int f = 0; f < 4; f++)
		sp[f] = cls[(4L * e + f) * kp.nv + v] == 1 ? flux_coef(ct, f, vv) : 0.0;
	double F[10][10];
	for (int i = 0; i < 10; i++)
		for (int j = 0; j < 10; j++)
			F[i][j] = -(a[0] * c_internal[i][j][0] + a[1] * c_internal[i][j][1] +
				    a[2] * c_internal[i][j][2]);
	for (int f = 0; f < 4; f++) {
		if (sp[f] == 0.0)
			continue;
		for (int p = 0; p < 6; p++)
			for (int q = 0; q < 6; q++)
				F[c_face2node[f][p]][c_face2node[f][q]] += sp[f] * c_face_scal[p][q];
	}
	for (int i = 0; i < 10; i++)
		for (int j = 0; j < 10; j++) {
			double s = 0.0;
			for (int k = 0; k < 10; k++)
				s += c_inv_mass[i][k] * F[k][j];
			D[i][j] = (i == j ? 1.0 : 0.0) + tau * s;
		}
	for (int k = 0; k < 3; k++)
		feat[k] = (float)(tau * a[k]);
	for (int f = 0; f < 4; f++)
		feat[3 + f] = (float)(tau * sp[f]);
}

__device__ inline double uniform_hash(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	z ^= z >> 31;
	return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* features x[k][7] and targets y[k][100] = D^-1 - I (the stored exact
 * inverse) of the items idx[k] */
__global__ void k_learned_data(const int2 *__restrict__ items, const int *__restrict__ idx,
			       int n, const double *__restrict__ geom,
			       const uint8_t *__restrict__ cls, KinPar kp, double dt,
			       const double *__restrict__ dinv, BatchTab bt,
			       float *__restrict__ x, float *__restrict__ y)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const int it = idx[k], v = items[it].x, e = items[it].y;
	double D[10][10];
	float f[7];
	item_block(geom, cls, kp, bt.ninst ? bt.dt[bt.einst[e]] : dt, v, e, D, f);
	for (int q = 0; q < 7; q++)
		x[(long)k * 7 + q] = f[q];
	for (int q = 0; q < 100; q++)
		y[(long)k * 100 + q] = (float)(dinv[100L * it + q] - (q / 10 == q % 10 ? 1.0 : 0.0));
}

/*
 * Block solves D u = r of the items idx[k] for 6 right-hand sides each
 * (uniform in [-1, 1], hashed from the item and the side) with P = I + pred
 * (the network's prediction, or the identity when pred is null): method 0
 * preconditioned Richardson u <- u + P (r - D u), u_0 = P r; method 1
 * right-preconditioned GMRES (gmres_block); to |r - D u| <= tol |r| within
 * maxit iterations (products with D), else the exact solution D^-1 r
 * (fallback). Per side: iterations (or -1 for a fallback) and
 * |u - D^-1 r| / |D^-1 r|.
 */
/* right-preconditioned GMRES on D P y = r, u = P y, restarted every
 * LG_RESTART products (Saad and Schultz; modified Gram-Schmidt, Givens
 * rotations), from u = 0 to |r - D u| <= tol |r| within maxit products;
 * returns the products made, or -1 */
#define LG_RESTART 10
__device__ inline int gmres_block(const double (&D)[10][10], const double (&P)[10][10],
				  const double (&r)[10], double rn2, double tol, int maxit,
				  double (&u)[10])
{
	for (int i = 0; i < 10; i++)
		u[i] = 0.0;
	const double lim = tol * sqrt(rn2);
	int total = 0;
	while (true) {
		double V[LG_RESTART + 1][10], H[LG_RESTART + 1][LG_RESTART], cs[LG_RESTART],
			sn[LG_RESTART], g[LG_RESTART + 1];
		double beta2 = 0.0;
		for (int i = 0; i < 10; i++) {
			double s = r[i];
			for (int j = 0; j < 10; j++)
				s -= D[i][j] * u[j];
			V[0][i] = s;
			beta2 += s * s;
		}
		const double beta = sqrt(beta2);
		if (beta <= lim)
			return total;
		if (total >= maxit || !(beta == beta))
			return -1;
		for (int i = 0; i < 10; i++)
			V[0][i] /= beta;
		for (int i = 0; i <= LG_RESTART; i++)
			g[i] = 0.0;
		g[0] = beta;
		int nc = 0;
		for (int j = 0; j < LG_RESTART && total < maxit; j++) {
			double z[10], w[10];
			for (int i = 0; i < 10; i++) {
				double s = 0.0;
				for (int l = 0; l < 10; l++)
					s += P[i][l] * V[j][l];
				z[i] = s;
			}
			for (int i = 0; i < 10; i++) {
				double s = 0.0;
				for (int l = 0; l < 10; l++)
					s += D[i][l] * z[l];
				w[i] = s;
			}
			total++;
			for (int q = 0; q <= j; q++) {
				double h = 0.0;
				for (int i = 0; i < 10; i++)
					h += w[i] * V[q][i];
				H[q][j] = h;
				for (int i = 0; i < 10; i++)
					w[i] -= h * V[q][i];
			}
			double hn = 0.0;
			for (int i = 0; i < 10; i++)
				hn += w[i] * w[i];
			hn = sqrt(hn);
			H[j + 1][j] = hn;
			for (int i = 0; i < 10; i++)
				V[j + 1][i] = hn > 0 ? w[i] / hn : 0.0;
			for (int q = 0; q < j; q++) {
				const double a0 = H[q][j], a1 = H[q + 1][j];
				H[q][j] = cs[q] * a0 + sn[q] * a1;
				H[q + 1][j] = -sn[q] * a0 + cs[q] * a1;
			}
			const double rho = sqrt(H[j][j] * H[j][j] + hn * hn);
			cs[j] = rho > 0 ? H[j][j] / rho : 1.0;
			sn[j] = rho > 0 ? hn / rho : 0.0;
			H[j][j] = rho;
			H[j + 1][j] = 0.0;
			g[j + 1] = -sn[j] * g[j];
			g[j] = cs[j] * g[j];
			nc = j + 1;
			if (fabs(g[j + 1]) <= lim || hn == 0.0)
				break;
		}
		double y[LG_RESTART];
		for (int q = nc - 1; q >= 0; q--) {
			double s = g[q];
			for (int l = q + 1; l < nc; l++)
				s -= H[q][l] * y[l];
			y[q] = s / H[q][q];
		}
		double x[10];
		for (int i = 0; i < 10; i++) {
			double s = 0.0;
			for (int q = 0; q < nc; q++)
				s += y[q] * V[q][i];
			x[i] = s;
		}
		for (int i = 0; i < 10; i++) {
			double s = 0.0;
			for (int l = 0; l < 10; l++)
				s += P[i][l] * x[l];
			u[i] += s;
		}
	}
}

__global__ void k_learned_solve(const int2 *__restrict__ items, const int *__restrict__ idx,
				int n, const double *__restrict__ geom,
				const uint8_t *__restrict__ cls, KinPar kp, double dt,
				const double *__restrict__ dinv, BatchTab bt,
				const float *__restrict__ pred, double tol, int maxit, int method,
				int *__restrict__ iters, double *__restrict__ relerr)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 6 * n)
		return;
	const int item = k / 6, side = k % 6;
	const int it = idx[item], v = items[it].x, e = items[it].y;
	double D[10][10], P[10][10];
	float f[7];
	item_block(geom, cls, kp, bt.ninst ? bt.dt[bt.einst[e]] : dt, v, e, D, f);
	for (int i = 0; i < 10; i++)
		for (int j = 0; j < 10; j++)
			P[i][j] = (i == j ? 1.0 : 0.0) +
				  (pred ? (double)pred[(long)item * 100 + 10 * i + j] : 0.0);
	double r[10], u[10], ex[10], rn = 0.0;
	for (int i = 0; i < 10; i++) {
		r[i] = 2.0 * uniform_hash(0x9e3779b97f4a7c15ull * (uint64_t)(7 * k + 1) + i) - 1.0;
		rn += r[i] * r[i];
	}
	for (int i = 0; i < 10; i++) {
		double s = 0.0;
		for (int j = 0; j < 10; j++)
			s += dinv[100L * it + 10 * i + j] * r[j];
		ex[i] = s;
		double t = 0.0;
		for (int j = 0; j < 10; j++)
			t += P[i][j] * r[j];
		u[i] = t;
	}
	int nit = -1;
	if (method == 1)
		nit = gmres_block(D, P, r, rn, tol, maxit, u);
	for (int m = 1; method == 0 && m <= maxit; m++) {
		double res[10], nr = 0.0;
		for (int i = 0; i < 10; i++) {
			double s = r[i];
			for (int j = 0; j < 10; j++)
				s -= D[i][j] * u[j];
			res[i] = s;
			nr += s * s;
		}
		if (!(nr <= tol * tol * rn) && !(nr == nr)) /* NaN: diverged */
			break;
		if (nr <= tol * tol * rn) {
			nit = m - 1;
			break;
		}
		for (int i = 0; i < 10; i++) {
			double s = 0.0;
			for (int j = 0; j < 10; j++)
				s += P[i][j] * res[j];
			u[i] += s;
		}
	}
	if (nit < 0)
		for (int i = 0; i < 10; i++)
			u[i] = ex[i]; /* exact fallback */
	double de = 0.0, ne = 0.0;
	for (int i = 0; i < 10; i++) {
		de += (u[i] - ex[i]) * (u[i] - ex[i]);
		ne += ex[i] * ex[i];
	}
	iters[k] = nit;
	relerr[k] = sqrt(de / (ne > 0 ? ne : 1.0));
}

static int *upload_idx(const int *idx, int n)
{
	int *d = (int *)dmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
	HIPCHECK(hipMemcpy(d, idx, sizeof(int) * n, hipMemcpyHostToDevice));
	return d;
}

/* training data of a kinetic context (after gdh_relax_setup): host arrays
 * x[n][7], y[n][100] for the items idx[0..n) (item = position in the sweep
 * schedule, 0 <= idx < nv * nelem) */
extern "C" void gdh_learned_data(gdh_ctx *c, int n, const int *idx, float *x, float *y)
{
	int *di = upload_idx(idx, n);
	float *dx = (float *)dmalloc(sizeof(float) * 7 * (size_t)n);
	float *dy = (float *)dmalloc(sizeof(float) * 100 * (size_t)n);
	k_learned_data<<<nblocks(n, 64), 64, 0, c->stream>>>(c->items, di, n, c->geom, c->cls,
							     c->kp, c->dt, c->dinv, c->bt, dx, dy);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(x, dx, sizeof(float) * 7 * n, hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipMemcpyAsync(y, dy, sizeof(float) * 100 * n, hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	for (void *q : { (void *)di, (void *)dx, (void *)dy })
		HIPCHECK(hipFree(q));
}

/*
 * Block solves of the items idx with the network (mlp) or, when mlp is
 * null, the identity as preconditioner, by Richardson (method 0) or GMRES
 * (method 1). out[0] mean iterations
 * of the solves that converged, out[1] largest, out[2] fallback rate,
 * out[3] largest relative L2 difference to the exact solution, out[4]
 * solves. iters (6 n, may be null): per solve, -1 = fallback.
 */
extern "C" void gdh_learned_solve_stats_ex(gdh_ctx *c, gdh_mlp *mlp, int n, const int *idx,
					   double tol, int maxit, int method, double *out,
					   int *iters)
{
	if (method != 0 && method != 1) {
		fprintf(stderr, "gdon3d-hip: block solve method %d (0 Richardson, 1 GMRES)\n",
			method);
		abort();
	}
	int *di = upload_idx(idx, n);
	float *pred = nullptr;
	if (mlp) {
		std::vector<float> hx(7 * (size_t)n), hy(100 * (size_t)n);
		gdh_learned_data(c, n, idx, hx.data(), hy.data());
		std::vector<float> hp(100 * (size_t)n);
		gdh_mlp_predict(mlp, hx.data(), n, hp.data());
		pred = (float *)dmalloc(sizeof(float) * 100 * (size_t)n);
		HIPCHECK(hipMemcpy(pred, hp.data(), sizeof(float) * 100 * n, hipMemcpyHostToDevice));
	}
	int *dit = (int *)dmalloc(sizeof(int) * 6 * (size_t)n);
	double *der = (double *)dmalloc(sizeof(double) * 6 * (size_t)n);
	k_learned_solve<<<nblocks(6L * n, 64), 64, 0, c->stream>>>(
		c->items, di, n, c->geom, c->cls, c->kp, c->dt, c->dinv, c->bt, pred, tol, maxit,
		method, dit, der);
	HIPCHECK(hipGetLastError());
	std::vector<int> hit(6 * (size_t)n);
	std::vector<double> her(6 * (size_t)n);
	HIPCHECK(hipMemcpyAsync(hit.data(), dit, sizeof(int) * hit.size(), hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipMemcpyAsync(her.data(), der, sizeof(double) * her.size(),
				hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	double sum = 0, mx = 0, err = 0;
	long conv = 0, fb = 0;
	for (size_t k = 0; k < hit.size(); k++) {
		if (hit[k] < 0)
			fb++;
		else {
			conv++;
			sum += hit[k];
			mx = std::max(mx, (double)hit[k]);
		}
		err = std::max(err, her[k]);
	}
	out[0] = conv ? sum / conv : 0.0;
	out[1] = mx;
	out[2] = (double)fb / (double)hit.size();
	out[3] = err;
	out[4] = (double)hit.size();
	if (iters)
		memcpy(iters, hit.data(), sizeof(int) * hit.size());
	for (void *q : { (void *)di, (void *)pred, (void *)dit, (void *)der })
		if (q)
			HIPCHECK(hipFree(q));
}

extern "C" void gdh_learned_solve_stats(gdh_ctx *c, gdh_mlp *mlp, int n, const int *idx,
					double tol, int maxit, double *out, int *iters)
{
	gdh_learned_solve_stats_ex(c, mlp, n, idx, tol, maxit, 0, out, iters);
}

/* ------------------------------------------------------------------ */
/* The network's block inverses inside the transport sweep               */
/* ------------------------------------------------------------------ */

/* Per sweep item: its block's numbers in binary64 (x[8 it + q]: tau a_k,
 * k < 3, then tau s_f on the outflow faces, 0 elsewhere; x[8 it + 7] = 0)
 * and the network's inputs in binary32 (xf[7 it + q]) */
__global__ void k_item_numbers(const int2 *__restrict__ items, int nitems,
			       const double *__restrict__ geom, const uint8_t *__restrict__ cls,
			       KinPar kp, double dt_, BatchTab bt, double *__restrict__ x,
			       float *__restrict__ xf)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	const int v = items[it].x, e = items[it].y;
	const double dt = bt.ninst ? bt.dt[bt.einst[e]] : dt_;
	double ct[9];
	for (int k = 0; k < 9; k++)
		ct[k] = geom[10L * e + k];
	const double tau = THETA_L * dt / geom[10L * e + 9];
	const double *vv = kp.vi + 3 * v;
	double q[8];
	for (int k = 0; k < 3; k++)
		q[k] = tau * (ct[0 + k] * vv[0] + ct[3 + k] * vv[1] + ct[6 + k] * vv[2]);
	for (int f = 0; f < 4; f++)
		q[3 + f] = cls[(4L * e + f) * kp.nv + v] == 1 ? tau * flux_coef(ct, f, vv) : 0.0;
	q[7] = 0.0;
	for (int k = 0; k < 8; k++)
		x[8L * it + k] = q[k];
	for (int k = 0; k < 7; k++)
		xf[7L * it + k] = (float)q[k];
}

/* the predictions (D^-1 - I, binary32) rounded to bfloat16 and packed two
 * per word: p[50 it + m] = entries 2 m (low half), 2 m + 1 (high half) */
__global__ void k_pack_pred(const float *__restrict__ pred, long n, unsigned *__restrict__ p)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= 50 * n)
		return;
	const unsigned lo = (unsigned short)to_bf16(pred[2 * k]);
	const unsigned hi = (unsigned short)to_bf16(pred[2 * k + 1]);
	p[k] = lo | (hi << 16);
}

/*
 * The sweep's refinement of each item, one thread per item and in the same
 * order of operations as the sweep's lanes (learned_inverse in sweep.hip):
 * steps[it] = S, the fewest steps (at least 2) after which the last
 * residual R = I - D X before the final update is at most tol (the result
 * then lies within rounding of the inverse), or 0 when LB_MAX_STEPS do not
 * reach it (the item keeps the stored exact inverse). dev[it] = max |X_S -
 * D^-1| / max |D^-1| against the stored exact inverse, res0[it] =
 * max |I - D X0| of the prediction.
 */
__global__ void k_learned_check(int nitems, const double *__restrict__ x,
				const unsigned *__restrict__ p, const double *__restrict__ dinv,
				double tol, int *__restrict__ steps, double *__restrict__ dev,
				double *__restrict__ res0)
{
	const int it = blockIdx.x * blockDim.x + threadIdx.x;
	if (it >= nitems)
		return;
	double xs[8], X[100], Dm[100], Rm[100];
	for (int k = 0; k < 8; k++)
		xs[k] = x[8L * it + k];
	for (int m = 0; m < 50; m++) {
		const unsigned w = p[50L * it + m];
		const int a = 2 * m, b = 2 * m + 1;
		X[a] = lb_pred((unsigned short)(w & 0xffffu)) + (a / 10 == a % 10 ? 1.0 : 0.0);
		X[b] = lb_pred((unsigned short)(w >> 16)) + (b / 10 == b % 10 ? 1.0 : 0.0);
	}
	for (int e = 0; e < 100; e++)
		Dm[e] = lb_entry(xs, e / 10, e % 10);
	int S = 0;
	for (int step = 0; step < LB_MAX_STEPS && !S; step++) {
		double res = 0.0;
		for (int e = 0; e < 100; e++) {
			double r = e / 10 == e % 10 ? 1.0 : 0.0;
			for (int l = 0; l < 10; l++)
				r -= Dm[10 * (e / 10) + l] * X[10 * l + e % 10];
			Rm[e] = r;
			res = fmax(res, fabs(r));
		}
		if (step == 0)
			res0[it] = res;
		double Xn[100];
		for (int e = 0; e < 100; e++) {
			double s = X[e];
			for (int l = 0; l < 10; l++)
				s += X[10 * (e / 10) + l] * Rm[10 * l + e % 10];
			Xn[e] = s;
		}
		for (int e = 0; e < 100; e++)
			X[e] = Xn[e];
		if (step >= 1 && res <= tol) /* NaN fails */
			S = step + 1;
	}
	double d = 0.0, m = 0.0;
	for (int e = 0; e < 100; e++) {
		d = fmax(d, fabs(X[e] - dinv[100L * it + e]));
		m = fmax(m, fabs(dinv[100L * it + e]));
	}
	steps[it] = S;
	dev[it] = d / m;
}

/*
 * Learned block inverses in the transport sweep (no counterpart in the
 * reference, whose KLU factors each block exactly): the network predicts
 * every item's D^-1 - I from its seven numbers, stored in bfloat16 (200 B
 * per item instead of the 800 B of the stored binary64 inverse); the sweep
 * rebuilds D from the numbers and refines the prediction by Newton-Schulz
 * steps to the binary64 inverse. Items whose refinement would not converge
 * (residual above tol before the last step) keep the stored exact inverse.
 * mlp null: every item back on the stored inverse.
 */
extern "C" void gdh_learned_enable(gdh_ctx *c, gdh_mlp *mlp, double tol, double *out)
{
	if (!c->scheme || !c->nitems) {
		fprintf(stderr, "gdon3d-hip: learned block inverses need a set-up kinetic context\n");
		abort();
	}
	const int n = c->nitems;
	out[0] = out[1] = out[2] = out[3] = 0.0;
	if (!mlp) {
		c->kp.lb_p = nullptr;
		c->kp.lb_x = nullptr;
		sweep_set_learned(c, nullptr);
		return;
	}
	if (!c->lb_x) {
		c->lb_x = (double *)dmalloc(sizeof(double) * 8 * (size_t)n);
		c->lb_p = (unsigned *)dmalloc(sizeof(unsigned) * 50 * (size_t)n);
	}
	float *xf = (float *)dmalloc(sizeof(float) * 7 * (size_t)n);
	float *pred = (float *)dmalloc(sizeof(float) * 100 * (size_t)n);
	int *ok = (int *)dmalloc(sizeof(int) * (size_t)n);
	double *dev = (double *)dmalloc(sizeof(double) * (size_t)n);
	double *r0 = (double *)dmalloc(sizeof(double) * (size_t)n);
	k_item_numbers<<<nblocks(n, 64), 64, 0, c->stream>>>(c->items, n, c->geom, c->cls, c->kp,
							      c->dt, c->bt, c->lb_x, xf);
	HIPCHECK(hipGetLastError());
	k_mlp_predict<<<((n + 15) / 16 + 3) / 4, 256, 0, c->stream>>>(mlp->w, mlp->wb, mlp->norm,
								       xf, n, pred);
	HIPCHECK(hipGetLastError());
	k_pack_pred<<<nblocks(50L * n, 256), 256, 0, c->stream>>>(pred, n, c->lb_p);
	HIPCHECK(hipGetLastError());
	k_learned_check<<<nblocks(n, 64), 64, 0, c->stream>>>(n, c->lb_x, c->lb_p, c->dinv, tol,
							       ok, dev, r0);
	HIPCHECK(hipGetLastError());
	std::vector<int> hok(n);
	std::vector<double> hdev(n), hr0(n);
	HIPCHECK(hipMemcpyAsync(hok.data(), ok, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipMemcpyAsync(hdev.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipMemcpyAsync(hr0.data(), r0, sizeof(double) * n, hipMemcpyDeviceToHost,
				c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	long good = 0;
	double dmax = 0.0;
	for (int k = 0; k < n; k++)
		if (hok[k]) {
			good++;
			dmax = std::max(dmax, hdev[k]);
		}
	c->kp.lb_p = c->lb_p;
	c->kp.lb_x = c->lb_x;
	sweep_set_learned(c, ok);
	HIPCHECK(hipStreamSynchronize(c->stream));
	for (void *q : { (void *)xf, (void *)pred, (void *)ok, (void *)dev, (void *)r0 })
		HIPCHECK(hipFree(q));
	std::nth_element(hr0.begin(), hr0.begin() + n / 2, hr0.end());
	out[0] = (double)good;
	out[1] = (double)(n - good);
	out[2] = dmax;
	out[3] = hr0[n / 2];
}
