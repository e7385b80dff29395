/*
 * dtvproj -- C ABI of the MI355X (gfx950) cone-beam projector pair.
 *
 * All array arguments are device pointers to C-contiguous FP32 data:
 *   volume   [batch][Z][Y][X]        (x fastest; the reference's [B,1,Z,Y,X])
 *   sinogram [batch][views][V][U]    (u fastest; the reference's [B,1,A,V,U])
 * Outputs are caller-allocated and fully overwritten; the library keeps no
 * pointer after a call returns and allocates no device memory.  Outputs
 * must be coarse-grained device memory (hipMalloc, as torch allocates):
 * some kernels add into them with hardware FP32 atomics, which are not
 * performed on managed, host or fine-grained memory, so the projector entry
 * points reject such outputs with DTV_ERR_INVALID_ARGUMENT.  Work is
 * enqueued on ``stream`` (a hipStream_t; NULL = the null stream) and runs
 * on the device that is current on the calling thread.  Every entry point
 * returns 0 on success or a DTV_ERR_* code; dtv_strerror() describes it.
 * Calls are reentrant: there is no global state.
 */
#ifndef DTVPROJ_H
#define DTVPROJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTV_OK 0
#define DTV_ERR_INVALID_ARGUMENT 1
#define DTV_ERR_NULL_POINTER 2
#define DTV_ERR_TOO_LARGE 3
#define DTV_ERR_HIP 4
#define DTV_ERR_VIEW_COUNT 5

/* Sample phase of a canonical view table (see dtv_cone_forward). */
typedef struct dtv_scheme {
  int32_t axis_mode;   /* 2: dominant axis among x,y (|dx|>=|dy| -> x); 3: among x,y,z (ties x->y->z) */
  int32_t reserved;
  double box_lo[3];    /* box whose dominant-axis planes phase the samples (voxel frame) */
  double box_hi[3];
  double margin;       /* voxels added at both ends of the sample range */
  double scale;        /* output multiplier */
} dtv_scheme;

/* Kernel choice for the *_ex entry points (0 fields = library defaults). */
typedef struct dtv_launch {
  int32_t algorithm;   /* 0: default, 1: ray-driven (gather / atomic scatter), 2: brick (LDS) */
  int32_t brick[3];    /* brick edge in voxels (x, y, z) for algorithm 2 */
  int32_t block;       /* threads per workgroup for algorithm 2, a multiple of 64: A 64..512
                          (default 512), A^T 64..256 (default 256) */
  int32_t view_parts;  /* algorithm 2: workgroups sharing one brick's views (0 = default;
                          A: 1..16; A^T: 1 or 2, with the output zero-filled and added to,
                          so it must be coarse-grained device memory) */
  int32_t reserved[2];
} dtv_launch;

#define DTV_ALGO_DEFAULT 0
#define DTV_ALGO_RAY 1
#define DTV_ALGO_BRICK 2

/* ---- Reference-shaped entry points ------------------------------------ */

/* Standard circular geometry, ray-driven forward projector A.
 * Replaces JITStandardGeometry.forward,
 * model/ConeBeamLayers/Standard/plug/include/forward.h:3 (kernel
 * Standard/plug/kernel/forwardKernel.cu:22-143).  angles: device float64
 * degrees [n_angles]; volume_size = {X, Y, Z}; detector_size = {U, V};
 * device: HIP ordinal the pointers live on (-1 = current device). */
int dtv_standard_forward(const float* volume, float* sino, int32_t batch,
                         const double* angles, int32_t n_angles,
                         const int32_t volume_size[3], const int32_t detector_size[2],
                         double sid, double sdd, int32_t device, void* stream);

/* Standard reference backprojector (voxel-driven, bilinear, FDK weights
 * (SDD/L)^2*pi/A; NOT the adjoint of A).  Replaces
 * JITStandardGeometry.backward, Standard/plug/include/backward.h:3 (kernel
 * Standard/plug/kernel/backwardKernel.cu:22-110). */
int dtv_standard_backward(const float* sino, float* volume, int32_t batch,
                          const double* angles, int32_t n_angles,
                          const int32_t volume_size[3], const int32_t detector_size[2],
                          double sid, double sdd, int32_t device, void* stream);

/* Standard matched adjoint: the exact transpose of dtv_standard_forward.
 * No reference counterpart (the reference wires backward.h:3 as A's
 * gradient, StandardGeometry.py:18-21); used for autograd and the
 * adjoint tests. */
int dtv_standard_adjoint(const float* sino, float* volume, int32_t batch,
                         const double* angles, int32_t n_angles,
                         const int32_t volume_size[3], const int32_t detector_size[2],
                         double sid, double sdd, int32_t device, void* stream);

/* Beijing vector geometry, forward projector A (without the Python-side
 * sampleInterval factor).  Replaces JITBeijingGeometry.forward,
 * model/ConeBeamLayers/Beijing/plug/include/forward.h:3 (kernel
 * Beijing/plug/kernel/forwardKernel.cu:17-135).  project_vector: device
 * float32 [n_views][12]. */
int dtv_beijing_forward(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const float* project_vector, int32_t n_views,
                        int32_t device, void* stream);

/* Beijing reference backprojector (triple-product projection, weight
 * fr*2pi/(n_views/21), skip rule on volume dimensions).  Replaces
 * JITBeijingGeometry.backward, Beijing/plug/include/backward.h:3 (kernel
 * Beijing/plug/kernel/backwardKernel.cu:18-104).  n_views must be a
 * multiple of 21 (the reference processes 21 views per launch). */
int dtv_beijing_backward(const float* sino, float* volume, int32_t batch,
                         const int32_t volume_size[3], const int32_t detector_size[2],
                         const float* project_vector, int32_t n_views,
                         int32_t device, void* stream);

/* Beijing matched adjoint: exact transpose of dtv_beijing_forward.  No
 * reference counterpart (BeijingGeometry.py:20-25 wires backward.h:3). */
int dtv_beijing_adjoint(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const float* project_vector, int32_t n_views,
                        int32_t device, void* stream);

/* ---- Canonical view-table entry points -------------------------------- */

/* Forward projector for any geometry given as a device float64 view table
 * [n_views][12] = S, D0, U, V in the voxel frame (voxel (ix,iy,iz) centred
 * at the integer point).  Generalises both reference kernels
 * (Standard forwardKernel.cu:22-86, Beijing forwardKernel.cu:17-76). */
int dtv_cone_forward(const float* volume, float* sino, int32_t batch,
                     const int32_t volume_size[3], const int32_t detector_size[2],
                     const double* views, int32_t n_views, const dtv_scheme* scheme,
                     void* stream);

/* Exact transpose of dtv_cone_forward. */
int dtv_cone_adjoint(const float* sino, float* volume, int32_t batch,
                     const int32_t volume_size[3], const int32_t detector_size[2],
                     const double* views, int32_t n_views, const dtv_scheme* scheme,
                     void* stream);

/* dtv_cone_forward / dtv_cone_adjoint with an explicit kernel choice (used
 * by the autotuner; launch = NULL means defaults).  Results of every
 * algorithm agree to FP32 rounding. */
int dtv_cone_forward_ex(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* stream);

/* Deterministic A: as dtv_cone_forward_ex, but each (ray, brick) partial sum
 * is added into an int64 fixed-point accumulator per ray (at a power-of-two
 * scale from the volume's largest |value|, so no sum can overflow) and
 * converted to FP32 at the end, so the result is bitwise reproducible for
 * any launch.  The workspace (device memory, 256-byte aligned) holds a
 * 256-byte header and 8 bytes per ray: dtv_cone_forward_plan gives the size.
 * A volume holding inf or NaN takes the FP32-atomic path (values propagate,
 * the last bits may vary).  No reference counterpart: the reference's
 * forward (forwardKernel.cu:115-141) launches per view and is deterministic
 * by construction; this restores that property for the brick kernels. */
int dtv_cone_forward_ws(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* workspace, size_t workspace_bytes,
                        void* stream);

/* The launch dtv_cone_forward_ws takes (NULL launch = defaults) and its
 * workspace bytes (0 for the ray-driven kernel).  No GPU work. */
int dtv_cone_forward_plan(int32_t batch, const int32_t volume_size[3],
                          const int32_t detector_size[2], int32_t n_views,
                          const dtv_launch* launch, dtv_launch* plan, size_t* workspace_bytes);
int dtv_cone_adjoint_ex(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* stream);

/* Matched A^T with the samples partitioned over the bricks: each sample is
 * marched by the one brick holding its floor cell (the dtv_cone_adjoint
 * bricks also march, and discard, the samples that only touch their halo,
 * about 17% more), a brick's taps on its upper-halo voxels are summed per
 * voxel and stored in the workspace, and a second kernel adds them into the
 * neighbouring bricks' voxels in one fixed order.  Each brick's views are
 * split over view_parts workgroups: one part stores into the volume, two add
 * into it (zero-filled by the call; FP32 addition of two terms does not depend
 * on their order), more store their partial volumes in the workspace and the
 * second kernel adds them in part order; so the result is bitwise
 * reproducible for any split (up to 16 parts).  The workspace (device memory,
 * dtv_cone_adjoint_plan gives its size) holds every part's halo sums and the
 * partial volumes of parts beyond the first.  A NULL workspace is taken for
 * one part (the unpartitioned kernel).  launch == NULL or view_parts == 0
 * takes the library's split for the volume size.  No reference counterpart:
 * it serves callers that can hand over scratch memory (the Python operators
 * pass a torch allocation); JITStandardGeometry.backward-shaped callers keep
 * dtv_*_adjoint. */
int dtv_cone_adjoint_ws(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, float* workspace, size_t workspace_bytes,
                        void* stream);

/* The launch dtv_cone_adjoint_ws takes for these sizes and `launch` (NULL =
 * defaults): kernel, brick, lanes and view parts in *plan, and the workspace
 * bytes it needs (4 x (parts x batch x bricks x halo voxels per brick +
 * (parts - 1) x batch x voxels), halo voxels per brick (bx+1)(by+1)(bz+1) -
 * bx by bz; 0 for the ray-driven kernel).  No GPU work. */
int dtv_cone_adjoint_plan(int32_t batch, const int32_t volume_size[3],
                          const int32_t detector_size[2], int32_t n_views,
                          const dtv_launch* launch, dtv_launch* plan, size_t* workspace_bytes);

/* ---- Uneven instances in one launch ----------------------------------- */

/* One volume/sinogram pair of a multi-instance call (canonical view table
 * geometry).  Forward: in = volume, out = sinogram (initialised to -init, or
 * 0 when init is NULL, then accumulated: out = A in - init).  Adjoint: in =
 * sinogram, out = volume (fully overwritten).  bx..bz is this instance's
 * brick shape, nbx..nbz its brick grid (forward bricks tile [-1, N-1]:
 * ceil((N+1)/b); adjoint bricks tile [0, N-1]: ceil(N/b)).  view_parts
 * (0 or 1: none) splits each brick's views over that many work items
 * (forward at most 16, adjoint at most 2; an adjoint instance with two parts
 * has its volume zero-filled by the call and added into, so it must be
 * coarse-grained device memory). */
typedef struct dtv_instance {
  const float* in;
  float* out;
  const float* init;
  const double* views;
  int32_t nx, ny, nz, nu, nv, n_views;
  int32_t axis_mode;
  int32_t bx, by, bz;
  int32_t nbx, nby, nbz;
  int32_t view_parts;
  double box_lo[3], box_hi[3];
  double margin, scale;
} dtv_instance;

/* One brick of one instance, or one part of its views: brick = b + nb * part
 * for brick b < nb = nbx * nby * nbz and part < the instance's view_parts. */
typedef struct dtv_work_item {
  int32_t instance;
  int32_t brick;
} dtv_work_item;

/* A (resp. matched A^T) over every (instance, brick) of ``work`` in one
 * kernel launch.  inst_host / work_host are host copies of the device
 * tables inst_dev / work_dev (validated and sized on the host; the kernel
 * reads the device copies, so the call can be captured in a graph).  With
 * counter != NULL (a device int) and persistent_blocks > 0, that many
 * workgroups (at most as many as the device holds at once for the kernel)
 * pull items from a queue (dynamic balance); otherwise one workgroup per
 * item, in list order.  block: lanes per workgroup, a multiple of 64, A up
 * to 512, A^T up to 256.  No reference counterpart: the
 * reference launches per view and per batch item (forwardKernel.cu:115-141). */
int dtv_cone_forward_multi(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                           int32_t n_inst, const dtv_work_item* work_host,
                           const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                           int32_t* counter, int32_t persistent_blocks, void* stream);
int dtv_cone_adjoint_multi(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                           int32_t n_inst, const dtv_work_item* work_host,
                           const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                           int32_t* counter, int32_t persistent_blocks, void* stream);

/* dtv_cone_adjoint_multi with the samples partitioned over the bricks (as
 * dtv_cone_adjoint_ws): each sample is marched by the one brick holding its
 * floor cell, each item stores its bricks' upper-halo sums in the workspace,
 * and a second kernel adds them into the neighbouring bricks' voxels in one
 * fixed order (bitwise reproducible; per instance equal to dtv_cone_adjoint_ws
 * with the same brick and view parts).  The workspace (device memory, 256-byte
 * aligned, dtv_cone_adjoint_multi_ws_bytes of the same host table) holds a
 * header of per-instance offsets, which the call computes on the device from
 * inst_dev (graph-capturable), and the halo sums.  No reference counterpart
 * (see dtv_cone_adjoint_multi). */
int dtv_cone_adjoint_multi_ws(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                              int32_t n_inst, const dtv_work_item* work_host,
                              const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                              int32_t* counter, int32_t persistent_blocks, void* workspace,
                              size_t workspace_bytes, void* stream);

/* Workspace bytes of dtv_cone_adjoint_multi_ws for this instance table
 * (256-byte header per 16 instances, then 4 x view parts x bricks x
 * ((bx+1)(by+1)(bz+1) - bx by bz) per instance).  No GPU work. */
int dtv_cone_adjoint_multi_ws_bytes(const dtv_instance* inst_host, int32_t n_inst,
                                    size_t* bytes);

/* ---- Fused reconstruction update ---------------------------------------- */

/* One instance of the fused FISTA / smoothed-TV step (device pointers). */
typedef struct dtv_tv_item {
  const float* y;      /* current (momentum) point */
  const float* g;      /* data-term gradient A^T (A y - b) */
  const float* xprev;  /* previous iterate */
  float* xnew;         /* out: [y - step (g + lam grad TV_eps(y))]_+ */
  float* ynew;         /* out: xnew + beta (xnew - xprev) */
  int32_t nx, ny, nz;
  int32_t nonneg;
  float step, lam, eps, beta;
} dtv_tv_item;

/* Fused step for ``batch`` equal-size volumes.  mode 0: Neumann forward
 * differences; mode 1: the reference's difference matrices
 * (model/PDHG/Utils/Gradients.py:8-47, incl. its D_y stride of Y).  xnew
 * may be xprev (in place); neither output may be y (the stencil reads y's
 * neighbours) and xnew != ynew, else DTV_ERR_INVALID_ARGUMENT.  The
 * reference has no fused counterpart: its unrolls apply torch.sparse
 * matrices per batch item (e.g. model/PDHG/ASDPOCS.py:38-43,
 * model/FISTA/DTVFISTA.py:29-36). */
int dtv_tv_fista_step(const float* y, const float* g, const float* xprev, float* xnew,
                      float* ynew, int32_t batch, const int32_t volume_size[3], float step,
                      float lam, float eps, float beta, int32_t nonneg, int32_t mode,
                      void* stream);

/* Same step over uneven instances in one launch.  items_dev is a device
 * array of n_items (<= 256) descriptors, items_host its host copy (read on
 * the host to size the launch: each instance gets exactly its own
 * workgroups; the kernel reads the device copy, so the call can be captured
 * in a graph). */
int dtv_tv_fista_step_multi(const dtv_tv_item* items_host, const dtv_tv_item* items_dev,
                            int32_t n_items, int32_t mode, void* stream);

/* One TV descent step of ASD-POCS (model/PDHG/ASDPOCS.py:37-43, whose
 * gradTV is built from the sparse matrices of Gradients.py:8-47):
 *   out = f - step * sum_a D_a^T (D_a f / (|D f| + eps))
 * with |D f| = sqrt(d_x^2 + d_y^2 + d_z^2) per voxel (the reference has
 * eps = 1).  f, out: device float32 [batch][Z][Y][X], distinct (the stencil
 * reads neighbours of f); mode as dtv_tv_fista_step; eps > 0. */
int dtv_tv_descent(const float* f, float* out, int32_t batch, const int32_t volume_size[3],
                   float step, float eps, int32_t mode, void* stream);

/* One primal-dual step of the learned directional-TV PDHG unroll, after the
 * caller has run A on fbar, updated w and run A^T on w into atw (the update
 * of model/PDHG/DTVCP.py:41-52, with the dual l1-ball projection of
 * L1ball.py:4-17 replaced by the box [-lam_a, lam_a]; D_a as in
 * dtv_tv_fista_step):
 *   p_a  <- clamp(p_a + sigma D_a fbar, -lam_a, lam_a)        (a = x, y, z)
 *   f'    = max(f - tau (atw + sum_a D_a^T p_a), 0)
 *   fbar <- 2 f' - f;  f <- f'
 * f, fbar, atw: device float32 [batch][Z][Y][X]; p: [3][batch][Z][Y][X]
 * (x, y, z planes); f, fbar and p are updated in place, all four distinct;
 * lam: host float[3] >= 0.  Two launches on ``stream``. */
int dtv_pdhg_dtv_step(float* f, float* fbar, const float* atw, float* p, int32_t batch,
                      const int32_t volume_size[3], float sigma, float tau, const float lam[3],
                      int32_t mode, void* stream);

/* ---- Fused shrinkage steps ----------------------------------------------- */

/* One vector of a batched l1-ball projection (device pointers). */
#define DTV_L1_MAX_ITEMS 16
typedef struct dtv_l1_item {
  const float* v;   /* input, n values */
  float* out;       /* output, n values (may be v) */
  int64_t n;
  double radius;    /* b */
  float sigma;      /* mode 1 only: the dual step size (> 0) */
  int32_t reserved;
} dtv_l1_item;

/* Workspace of dtv_l1ball for n_items vectors (0 when n_items is out of
 * range); device memory, 8-byte aligned, no initialisation needed. */
size_t dtv_l1ball_workspace_bytes(int32_t n_items);

/* l1-ball projection of up to DTV_L1_MAX_ITEMS vectors in one call,
 * following model/PDHG/Utils/L1ball.py:4-17: with a_i = |v_i| (mode 0) or
 * |v_i| / sigma (mode 1), if sum a_i < b the ball holds the point;
 * otherwise lam solves sum relu(a_i - lam) = b by the reference's Newton
 * iteration from lam = 0 (slope 0.001 - #{a_i > lam}, stop once
 * |sum relu(a_i - lam) - b| <= tol or after max_iter steps; the reference
 * has tol = 1e-4, max_iter = 100).  Writes
 *   mode 0: out = v if inside, else sign(v) relu(|v| - lam)
 *   mode 1: out = v - sigma sign(v) (a if inside, else relu(a - lam)),
 *           the dual update p = p' - sigma sign(p') l1ball(|p'| / sigma, b)
 *           of model/PDHG/DTVCP.py:44-46.
 * Runs entirely on ``stream`` with no host synchronisation (one launch per
 * Newton step; steps after convergence return at once).  No reference
 * counterpart on the device: the reference loops in Python with a host
 * synchronisation per step. */
int dtv_l1ball(const dtv_l1_item* items, int32_t n_items, int32_t mode, double tol,
               int32_t max_iter, void* workspace, size_t workspace_bytes, void* stream);

/* dtv_l1ball without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  No counterpart in the reference.  The state and
 * partial sums live in a static device array of 131,584 bytes (room for
 * DTV_L1_MAX_ITEMS vectors of any length), one per device, held until the
 * process ends; the call allocates nothing.  n_items > DTV_L1_MAX_ITEMS
 * returns DTV_ERR_TOO_LARGE.  Every call on one device shares that array
 * (each call clears the state it needs first): calls that overlap on one
 * device, e.g. on two streams, race; calls on one stream do not. */
int dtv_l1ball_noalloc(const dtv_l1_item* items, int32_t n_items, int32_t mode, double tol,
                       int32_t max_iter, void* stream);

/* Soft threshold with FISTA momentum over n values (model/FISTA/FISTA.py:21-25):
 *   d = y - step * g;  xnew = sign(d) relu(|d| - thr);  ynew = xnew + beta (xnew - xprev)
 * xnew may be xprev; ynew must differ from xprev and from xnew. */
int dtv_shrink_step(const float* y, const float* g, const float* xprev, float* xnew,
                    float* ynew, int64_t n, float step, float thr, float beta, void* stream);

/* Workspace of dtv_fdpg_step (device memory, 4-byte aligned). */
size_t dtv_fdpg_workspace_bytes(void);

/* One step of the directional-TV dual-gradient unroll of
 * model/FISTA/DTVFISTA.py:29-41 after the caller's g = A^T (A f - sino):
 *   z = relu(f - sigma g);  m_a = max over all voxels of the batch of |D_a z| - v_a;
 *   p_a = sign(D_a z) m_a;  fnew = z + sum_a t_a D_a^T p_a;
 *   fout = fnew + beta (fnew - f)
 * (D_a as in dtv_tv_fista_step, mode 0 Neumann / 1 reference matrices).
 * f, g, fout: [batch][Z][Y][X], fout distinct from f and g.  Three launches. */
int dtv_fdpg_step(const float* f, const float* g, float* fout, int32_t batch,
                  const int32_t volume_size[3], float sigma, const float v[3], const float t[3],
                  float beta, int32_t mode, void* workspace, size_t workspace_bytes,
                  void* stream);

/* ---- FBP front end ------------------------------------------------------ */

/* Cosine weighting fused with the per-row filtering of the FBP modules:
 *   output[r][x] = sum_k taps[k] * weight[r % weight_rows][j] * input[r][j],
 *   j = x + k - (n_taps - 1) / 2, terms with j outside [0, width) dropped,
 * i.e. ``conv2d(input * cosWeight, ramp, padding=(0, width/2))`` of
 * StandardGeometry.py:55-65 and BeijingGeometry.py:60-63 (a cross-
 * correlation; n_taps odd, output width = width).  input/output: device
 * float32 [rows][width], distinct; weight: device [weight_rows][width] or
 * NULL for none; taps: device float32 [n_taps].  Needs
 * (width + 2*roundup8(n_taps) + 8) * 4 bytes <= 64 KiB. */
int dtv_row_filter(const float* input, float* output, int64_t rows, int32_t width,
                   const float* weight, int32_t weight_rows, const float* taps,
                   int32_t n_taps, void* stream);

/* ---- Host-only helpers (no GPU work) ---------------------------------- */

/* Canonical view table and scheme of the Standard geometry, computed on the
 * host with the same code the kernels run (views_out: host [n][12]). */
int dtv_views_standard(const double* angles, int32_t n_angles,
                       const int32_t volume_size[3], const int32_t detector_size[2],
                       double sid, double sdd, double* views_out, dtv_scheme* scheme_out);

/* Same for Beijing projection vectors (host float32 [n][12]). */
int dtv_views_beijing(const float* project_vector, int32_t n_views,
                      const int32_t volume_size[3], const int32_t detector_size[2],
                      double* views_out, dtv_scheme* scheme_out);

const char* dtv_strerror(int code);
/* "dtvproj <version> gfx950 src:<16 hex digits>": the digits are the first of
 * the SHA-256 over the sources the library was compiled from (its build
 * script computes them), so a loader can refuse a stale build. */
const char* dtv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DTVPROJ_H */
