/* C-ABI of libmr_recon_amd.so -- MI355X (gfx950) kernels of the mr_recon NUFFT hot path.
 *
 * Every entry point takes raw device pointers, sizes and the HIP stream to
 * launch on (void*, a hipStream_t); it enqueues work and returns without
 * synchronising. Return value: 0 on success, non-zero on a validation or launch
 * error, with the message available from mra_last_error(). The library never
 * allocates or frees device memory: callers pass every buffer, including
 * workspaces, whose sizes the *_sizes / *_workspace / *_points queries give.
 *
 * complex64 buffers are interleaved (re, im) float pairs, the layout of
 * torch.complex64. References "file:line" are to danielabrahamgit/mr_recon.
 */
#ifndef MR_RECON_AMD_H
#define MR_RECON_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------------------- */
const char* mra_last_error(void);
int mra_version(void);

/* ---- geometry records ------------------------------------------------------------------
 * A Geom record describes one problem (grid size, batch, sample count, kernel) for
 * the interp / gridding / gradient launches; a table of them in device memory lets
 * one launch serve many uneven problems. */
int mra_geom_size(void);
int mra_default_tile(int ndim);
/* Tile edge (grid points) of the work decomposition. The kernels are compiled for 2D tiles
 * of 8, 16 (default) and 32 and for one edge in 1D (256) and 3D (8). mra_set_tile selects the
 * edge that the calling thread's subsequent plan, Geom, subgrid-size, coil-chunk and kernel
 * calls for ndim use (returns the previous edge, or -1 for an edge without kernels); a plan
 * and every call on it must use the same edge. No counterpart in the reference (its Triton
 * kernels have no tiles); it exists so that an autotuner can choose the tile. */
int mra_tile(int ndim);
int mra_set_tile(int ndim, int tile);
long long mra_subgrid_points(int ndim, int P);
int mra_max_coil_chunk(int ndim, int P);
int mra_grad_samples_per_block(void);
int mra_geom_fill(void* dst_host, long long grid_off, long long grid_stride, long long ksp_off,
                  long long ksp_stride, long long samp_off, int nbatch, int npts, int ndim, const int* G,
                  int tile, const int* wrap, float beta, float width, float out_scale, int item_base,
                  long long scratch_off, int tile_items_off);

/* ---- sample plans (binning) -------------------------------------------------------------
 * Replaces the per-call coordinate handling of the reference:
 *   triton_nufft._scale_trj          fourier.py:420-430   t = fl32(fl32(k*fl32(G/N)) + G//2)
 *   prep_ungrid_shapes / prep_grid   ungrid.py:851-894, grid.py:508-553   t = remainder(t, G)
 *   get_neighborhood                 triton_interp/kernels.py:531-536     lower = ceil(t - w/2)
 * bit for bit, then sorts samples stably by (tile of their patch's lower corner, position
 * of that corner in the tile) with LSD radix passes. The work area holds
 * mra_bin_workspace(...) 4-byte words; tile_start gets ntiles + 2 entries (tile ntiles
 * collects non-finite samples, the last entry is npts). */
long long mra_scan_workspace(long long n);
int mra_scan_excl_i32(const int* in, long long n, int* out, int* ws, void* stream);
long long mra_bin_workspace(int ndim, long long npts, long long ntiles, int tile, int chunk_len);
int mra_bin_build(int ndim, const float* coords, long long npts, const int* G, const float* scale,
                  const float* shift, int apply_scale, float width, int tile, const int* wrap, int chunk_len,
                  void* work, float* tsort, int* perm, int* tile_start, void* stream);
int mra_bin_items(const int* tile_start, int ntiles, int S, int prob, int* nitems_ws, int* item_off,
                  int* scan_ws, void* items, int max_items, void* stream);
/* Kaiser-Bessel weights of every sorted sample (kernels.py:180-246 kaiser_bessel,
 * kernels.py:295-393 weights2d/weights3d, norm '1'). */
int mra_plan_weights(int ndim, int P, const float* tsort, int npts, const int* G, const int* wrap, float width,
                     float beta, float* wsort, int* lbase, void* stream);
/* Same with one Kaiser-Bessel beta per dimension (betas[ndim]): sigpy's interpolate / gridding
 * with a per-dimension `param`, as type3_nufft grids on (linops.py:250-251, 311-312). */
int mra_plan_weights_dims(int ndim, int P, const float* tsort, int npts, const int* G, const int* wrap, float width,
                          const float* betas, float* wsort, int* lbase, void* stream);

/* ---- interpolation / gridding -----------------------------------------------------------
 * mra_interp replaces ungrid (ungrid.py:34-124, Triton kernels _ungrid1d/2d/3d
 * ungrid.py:146-495) for kernel='kaiser_bessel', norm='1', pad_mode='circular'. */
int mra_interp(int ndim, int P, const void* grid, void* out, const float* wsort, const int* lbase,
               const int* perm, const void* geoms, const void* items, int nitems, int ccoils, int nchunks,
               void* stream);
/* mra_grid_partial + mra_grid_reduce replace grid (grid.py:32-128, Triton kernels
 * _grid1d/2d/3d grid.py:150-502, which scatter with fp32 atomics). No atomics
 * are used: phase 1 accumulates per tile in LDS with plain read-modify-writes and
 * stores subgrids to scratch; the reduce writes every grid point once, so the output
 * needs no zeroing. */
int mra_grid_partial(int ndim, int P, const void* ksp, const float* dcf, const float* wsort, const int* lbase,
                     const int* perm, const void* geoms, const void* items, int nitems, int nbatch,
                     void* scratch, void* stream);
int mra_grid_reduce(int ndim, int P, void* grid, const void* scratch, const void* geoms, int prob0, int nprob,
                    const int* tile_items, int ntiles, int nbatch, int accumulate, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): mra_grid_partial + mra_grid_reduce of one
 * problem whose phase-1 subgrids live in a static 64 MiB device array of the library (one
 * per GPU, held until the process ends) instead of a caller buffer. `geom_host` is the
 * problem's host Geom record (mra_geom_fill, item_base = scratch_off = tile_items_off = 0);
 * the call copies it to the device with a kernel and allocates nothing. Fails (status 1,
 * message in mra_last_error) when nitems x nbatch x mra_subgrid_points(ndim, P) x 8 bytes
 * exceeds mra_grid_noalloc_max_bytes(). The static scratch is shared by every call on one
 * GPU: calls that overlap on one GPU (on different streams, or from several host threads)
 * race and must be serialised by the caller. The stream must belong to the current device. */
int mra_grid_noalloc(int ndim, int P, const void* ksp, const float* dcf, const float* wsort, const int* lbase,
                     const int* perm, const void* geom_host, const void* items, int nitems, int nbatch,
                     const int* tile_items, int ntiles, void* grid, int accumulate, void* stream);
long long mra_grid_noalloc_max_bytes(void);
/* Marks a host Geom record (from mra_geom_fill) so that the gridding kernels index dcf by
 * sorted (plan) position, dcf_sorted[s] = dcf[perm[s]], instead of through perm. No
 * reference counterpart (the reference multiplies ksp by dcf before gridding,
 * linops.py:1208). */
int mra_geom_set_dcf_plan_order(void* geom_host, int on);
/* Interpolation output, gridding input and dcf all in sorted (plan) order: position s of a
 * problem's sample vectors is its s-th sorted sample (no perm gather or scatter). Used by
 * density_compensation (algs.py:20-106), whose iteration never needs sample order. */
int mra_geom_set_ksp_plan_order(void* geom_host, int on);
/* Selects the phase-1 kernel behind mra_grid_partial: 1 (default) = 3D cells of samples
 * summed on the matrix cores (P <= 7), 0 = per-lane LDS read-modify-write. Same scratch
 * output; returns the previous selection. No reference counterpart (one kernel per
 * dimensionality in grid.py:150-502). */
int mra_set_grid_kernel(int mode);
/* Gradient of Re<a, interp(grid, t)> with respect to the sample coordinates t.
 * No reference counterpart: InterpolateFn.backward returns None for locs
 * (triton_interp/interp.py:66-74). */
int mra_interp_grad(int ndim, int P, const void* grid, const void* a, float* grad_t, const float* tsort,
                    const int* perm, const void* geoms, const void* items, int nitems, int max_item_samples,
                    void* stream);

/* ---- around the FFT ---------------------------------------------------------------------
 * mra_pad_expand replaces, fused: sense_linop coil expand mps*img (linops.py:1161),
 * apodize_img (fourier.py:398-418), the /sqrt(prod N) scale and resize zero-pad
 * (fourier.py:456-462, utils.py:41-84) and the ifftshift of fft (fourier.py:45),
 * or PadLast.forward (pad.py:28-31) for the Toeplitz path. */
int mra_pad_expand(int ndim, const int* N, const int* G, const int* center, const void* x, long long nx,
                   long long x_stride, const void* smaps, long long ns, long long s_stride, const float* apod,
                   float scale, void* out, void* stream);
/* mra_crop_reduce replaces, fused: the fftshift of ifft (fourier.py:58), resize crop
 * and scaling (fourier.py:520-521), apodize_img, and the coil reduction
 * einsum(FHWy, conj(mps)) (linops.py:1220) / PadLast.adjoint (pad.py:33-37). */
int mra_crop_reduce(int ndim, const int* N, const int* G, const int* center, const void* grid, long long nx,
                    const void* smaps, long long ns, long long s_stride, int conj_maps, const float* apod,
                    float scale, int accumulate, int roll_out, void* out, long long x_stride, void* stream);
/* Batched problems of one image size (no reference counterpart: the reference loops
 * over problems in Python). Grid q pads image img_index[q] times map q; image i reduces
 * grids/maps coil_off[i] .. coil_off[i+1]-1. */
int mra_pad_expand_batched(int ndim, const int* N, const int* G, const int* center, const void* x, long long nimg,
                           const void* smaps, long long nq, const int* img_index, const float* apod, float scale,
                           void* out, void* stream);
int mra_crop_reduce_batched(int ndim, const int* N, const int* G, const int* center, const void* grid,
                            long long nimg, const void* smaps, const int* coil_off, const float* apod, float scale,
                            void* out, void* stream);
/* Toeplitz PSF multiply (fourier.py:344, linops.py:1285-1287). */
int mra_cmul_bcast(void* data, const void* k, long long nbatch, long long n, void* stream);
/* In-place unnormalised batched C2C FFT over the trailing ndim dims (rocFFT via hipFFT);
 * replaces the centred fft / ifft of the NUFFT (fourier.py:37-61, 489-525), whose shifts
 * are index maps of the pad / crop / KB kernels here. */
int mra_fft_plan(int ndim, const int* n, int batch, void** plan);
int mra_fft_exec(void* plan, void* data, int inverse, void* stream);
int mra_fft_destroy(void* plan);
/* The same transforms with one LDS-staged Stockham pass per axis (lengths whose radix plan
 * is compiled in: 64, 80, 96, 120, 128, 160, 192, 240, 256, 320, 384, 480, 512, 640; every
 * axis but the last needs >= 64 points after it). tw[d] is the device table of
 * exp(-2 pi i m / n[d]), m < n[d], filled on the host by mra_fft_lds_twiddles. */
int mra_fft_lds_supported(int ndim, const int* n);
int mra_fft_lds_twiddles(int L, float* tw_host);
int mra_fft_lds(int ndim, const int* n, long long batch, void* data, const void* const* tw, int inverse,
                void* stream);
/* FFTs fused with the pad / crop around them, skipping the known zeros of the padded
 * grid (fourier.py:432-468 forward_FT_only, 505-525 adjoint_iFT_only, 302-352
 * normal_toeplitz; SENSE expand / reduce linops.py:1161, 1220, 1265-1306).
 * mra_fft_pad_forward: grid[q] = FFT(pad(x[img(q)] * S[map(q)] * apod * scale));
 * mra_fft_crop_adjoint: out[i] (+)= scale * apod * sum_q crop(iFFT(grid[q])) * (conj) S[q]
 * (the grids are overwritten); mra_toeplitz_fused: the PSF normal operator of one image,
 * work = ns grids. Same grid shapes as mra_fft_lds. */
int mra_fft_fused_supported(int ndim, const int* G);
int mra_fft_pad_forward(int ndim, const int* N, const int* G, const int* center, const void* x, long long x_stride,
                        const void* smaps, long long ns, long long s_stride, const int* img_index, long long nq,
                        const float* apod, float scale, void* grid, const void* const* tw, void* stream);
int mra_fft_crop_adjoint(int ndim, const int* N, const int* G, const int* center, void* grid, long long nq,
                         long long nx, const void* smaps, long long ns, long long s_stride, const int* coil_off,
                         int conj_maps, const float* apod, float scale, int accumulate, void* out,
                         long long x_stride, const void* const* tw, void* stream);
int mra_toeplitz_fused(int ndim, const int* N, const int* G, const int* center, const void* x, const void* smaps,
                       long long ns, long long s_stride, const void* kern, void* work, int accumulate, void* out,
                       const void* const* tw, void* stream);

/* ---- nearest-grid sampling ------------------------------------------------------------------
 * Gather / scatter-add at one flat grid index per sample: gridded_nufft's forward_interp_only
 * and adjoint_grid_only (fourier.py:1250-1294, via indexing.multi_index / multi_grid,
 * indexing.py:29-79). idx values must lie in [0, gstride); the scatter adds into grid. */
int mra_nn_gather(const void* grid, long long gstride, const int* idx, long long M, int B, float scale, void* out,
                  void* stream);
int mra_nn_scatter(const void* ksp, const int* idx, long long M, int B, float scale, void* grid, long long gstride,
                   void* stream);
/* The same sums over a stable sort by flat index (perm: sorted -> sample), cut into chunks of
 * one grid point's samples (chunk c = sorted positions cstart[c] .. cend[c]-1 at point cidx[c]):
 * each chunk is summed in sorted order and added with one atomic per value. */
int mra_nn_scatter_chunks(const void* ksp, const int* perm, const int* cstart, const int* cend, const int* cidx,
                          long long NC, long long M, int B, float scale, void* grid, long long gstride,
                          void* stream);
/* Low-rank forms of the nearest-grid transform, chebyshev_nufft.forward / adjoint
 * (fourier.py:1703-1774: multi_index of L spectra times the temporal functions h, summed
 * over L; multi_grid of k-space times conj(h)). grid holds B * L grids of gstride points,
 * h is (L, M) complex64. gather_sum: out[b, m] = scale * sum_l h[l, m] grid[b L + l, idx[m]].
 * scatter_sum_chunks: adds scale * sum_{s in chunk c} conj(h[l, s]) ksp[b, s] into
 * grid[b L + l, cidx[c]] (chunk tables as for mra_nn_scatter_chunks). */
int mra_nn_gather_sum(const void* grid, long long gstride, const int* idx, const void* h, int L, long long M, int B,
                      float scale, void* out, void* stream);
int mra_nn_scatter_sum_chunks(const void* ksp, const void* h, int L, const int* perm, const int* cstart,
                              const int* cend, const int* cidx, long long NC, long long M, int B, float scale,
                              void* grid, long long gstride, void* stream);

/* ---- 4D and 5D interpolation / gridding ------------------------------------------------------
 * ungrid / grid for 4- and 5-dimensional grids: the Triton kernels _ungrid4d / _ungrid5d
 * (triton_interp/ungrid.py:498-845) and grid's torch fallback there (grid.py:86-127, 556-624).
 * t: M x ndim sample coordinates in grid units, already wrapped to [0, G] (prep_ungrid_shapes);
 * G, betas: ndim entries; P = floor(width)+1 in 2..7; every weight times out_scale. interp:
 * out[b, m] = sum over the patch of KB weight x grid[b, g] (B grids of prod(G) points); grid:
 * adds the samples into grid with float atomics (the caller zeroes it). */
int mra_interp_direct(int ndim, int P, const void* grid, const float* t, long long M, int B, const int* G,
                      const float* betas, float width, float out_scale, void* out, void* stream);
int mra_grid_direct(int ndim, int P, const void* ksp, const float* t, long long M, int B, const int* G,
                    const float* betas, float width, float out_scale, void* grid, void* stream);

/* ---- low-rank coil / imperfection mixing ---------------------------------------------------
 * imperf_coil_lowrank.forward / adjoint (linops.py:614-714): the einsums 'L ..., L C ... -> C ...'
 * and 'C ..., L C ... -> L ...' with conj(h) and the dcf, h = temporal functions (L, C, M)
 * complex64. expand: out[c, m] = sum_l h[l, c, m] y[l, m]; reduce: out[l, m] = dcf[m] *
 * sum_c conj(h[l, c, m]) ksp[c, m] (dcf may be NULL). */
int mra_lowrank_expand(const void* y, const void* h, int L, int C, long long M, void* out, void* stream);
int mra_lowrank_reduce(const void* ksp, const void* h, const float* dcf, int L, int C, long long M, void* out,
                       void* stream);

/* ---- density compensation ---------------------------------------------------------------
 * Elementwise steps of the Pipe-Menon iteration behind density_compensation (algs.py:20-106,
 * which calls sigpy.mri.pipe_menon_dcf): w /= |G^H G w| after the gridding and interpolation
 * calls above, and the final dcf /= max(dcf) (algs.py:105), optionally stored through the plan's
 * sort permutation (out[perm[i]]) when the iteration ran in plan order. part: mra_dcf_partials()
 * floats. */
int mra_dcf_partials(void);
int mra_dcf_update(const void* ghg, float* w, long long n, void* stream);
int mra_dcf_normalize(const float* w, const int* perm, float* out, long long n, float* part, void* stream);

/* ---- conjugate gradient -----------------------------------------------------------------
 * Vector work of conjugate_gradient (algs.py:779-872) with FP64 inner products
 * and device-resident step sizes. */
int mra_cg_partials(void);
int mra_cdot_re(const void* a, const void* b, long long n, float lam, double* part, double* sc, int slot,
                void* stream);
int mra_cg_init(const void* b, const void* ax, const void* x, void* r, void* p, long long n, float lam,
                double* part, double* sc, void* stream);
int mra_cg_step(void* x, void* r, void* p, const void* ap, long long n, float lam, double* part, double* sc,
                void* stream);
/* One CG step as a differentiable function, for unrolling CG under autograd: forward out of
 * place from nrm = (A^H A) p: ap = inv * nrm + lam * p, pAp, alpha, x1, r1, rz1 = <r1, r1>,
 * beta, p1 (FP64 dots; sd = double[8]: rz, pAp, rz1, beta, alpha), and the adjoint of that
 * step for the cotangents of (x1, r1, p1, rz1), any of which may be null. The reference has
 * no fused counterpart: its CG (algs.py:779-872) is differentiated op by op by torch
 * autograd. part: mra_cg_partials() doubles (forward), 4 x that (backward). */
int mra_cg_step_fwd(const void* x, const void* r, const void* p, const void* nrm, const double* rz, long long n,
                    float inv, float lam, void* x1, void* r1, void* p1, double* rz1, double* sd, double* part,
                    void* stream);
int mra_cg_step_bwd(const void* p, const void* nrm, const void* r1, const double* sd, const void* gx1,
                    const void* gr1, const void* gp1, const double* grz1, long long n, float inv, float lam,
                    void* gr, void* gp, void* gnrm, double* grz, double* part, double* bs, void* stream);
/* nseg independent systems of seg_len elements (batched problems; the reference runs
 * one conjugate_gradient per problem). */
int mra_cg_seg_blocks(void);
int mra_cg_init_seg(const void* b, const void* ax, const void* x, void* r, void* p, long long seg_len, int nseg,
                    float lam, double* part, double* sc, void* stream);
int mra_cg_step_seg(void* x, void* r, void* p, const void* ap, long long seg_len, int nseg, float lam,
                    double* part, double* sc, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MR_RECON_AMD_H */
