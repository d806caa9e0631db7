/* opem.h - C ABI of the MI355X per-frequency solve path (OpenParEM3D drop-in for one path).
 *
 * One context = one GPU + one HIP stream.  Caller owns every host buffer; the context owns all
 * device memory.  Every call returns 0 on success, a negative value for misuse (bad argument,
 * wrong call order) and a positive value for a runtime failure (HIP error, allocation);
 * opem_last_error() gives the message.  Nonconvergence of a Krylov solve is NOT an error: it
 * is reported through per-system `converged` flags, as the reference does
 * (src/OpenParEM3D/solveComplexLinearSystem.c:204-206, *converged).
 *
 * Complex numbers are interleaved (re, im) doubles.  Boundary attributes and volume attributes are
 * 1-based, as in MFEM meshes.  No exceptions cross the ABI.  One host thread per context.
 */
#ifndef OPEM_H
#define OPEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct opem_ctx opem_ctx;

/* Boundary kinds (src/OpenParEM3D/port.cpp:733-755 Boundary types; ports from Port definitions). */
enum {
  OPEM_BDR_PEC = 1,        /* essential E_t = 0 (fem3D.cpp:80-99, eliminatePEC)                     */
  OPEM_BDR_PMC = 2,        /* natural, nothing assembled                                            */
  OPEM_BDR_PORT = 3,       /* wave port: Dirichlet when driven, 1st-order 2D ABC otherwise           */
  OPEM_BDR_IMPEDANCE = 4,  /* surface impedance; value = conductivity sigma [S/m] (R_s from sigma)   */
  OPEM_BDR_RADIATION = 5   /* 1st-order radiation; value = wave impedance eta [ohm] (0 -> eta0)      */
};

enum { OPEM_PREC_FP64 = 0, OPEM_PREC_MIXED = 1 };

/* Context lifecycle. */
int opem_create(opem_ctx **ctx, int device);
int opem_destroy(opem_ctx *ctx);
const char *opem_last_error(const opem_ctx *ctx);

/* Mesh (replaces the reference's gmsh/MFEM mesh load, src/OpenParEMCommon/mesh.cpp:386-449).
 * xyz: nv*3, tets: nt*4 (0-based vertex ids), tet_attr: nt, bfaces: nbf*3, bface_attr: nbf. */
int opem_mesh_set(opem_ctx *ctx, int64_t nv, const double *xyz, int64_t nt, const int32_t *tets,
                  const int32_t *tet_attr, int64_t nbf, const int32_t *bfaces, const int32_t *bface_attr);

/* ND_p space, DOF numbering, shared CSR pattern, and the unit-coefficient base matrices
 * (curl-curl, mass, mixed curl per volume attribute) assembled on the GPU.
 * Replaces fem3D::build_fe_spaces (fem3D.cpp:65) + the MFEM assembly inside build_A/build_P/build_Q
 * (fem3D.cpp:103-225).  Materials and boundaries must be set before opem_sweep, not before this. */
int opem_space_build(opem_ctx *ctx, int order, int64_t *ndof, int64_t *nnz);

/* Read back the numbering / pattern (any pointer may be NULL).  elem_dofs: nt*nd, bface_dofs:
 * nbf*ndf, rowptr: ndof+1, colind: nnz. */
int opem_space_get(opem_ctx *ctx, int32_t *elem_dofs, int32_t *bface_dofs, int32_t *rowptr, int32_t *colind);

/* Unit-coefficient base volume matrix values on the pattern for one volume attribute.
 * which: 0 = curl-curl K, 1 = mass M, 2 = mixed curl Q (Q[a][b] = int curl(phi_b).phi_a). out: nnz. */
int opem_base_matrix(opem_ctx *ctx, int which, int attr, double *out);

/* Material of a volume attribute: eps = er*eps0*(1 - j tand), mu = mur*mu0
 * (src/OpenParEMCommon/OpenParEMmaterials.cpp:600-708). */
int opem_material_set(opem_ctx *ctx, int attr, double er, double tand, double mur);

/* Boundary attribute kind and parameter (see OPEM_BDR_*). */
int opem_boundary_set(opem_ctx *ctx, int attr, int kind, double value);

/* Quadrature points of a port at which the caller samples its modal fields (npts*3 doubles).
 * Call with xyz == NULL to query npts. */
int opem_port_points(opem_ctx *ctx, int attr, int64_t *npts, double *xyz);

/* Complex system matrix A(omega) values (nnz complex) with the 2D ABC on every port
 * (gamma: one complex per port, ports in ascending attribute order).  Mirrors fem3D::build_A
 * (fem3D.cpp:103-190) + hypre_ParCSRMatrixToMat (csr.c:171-223), before any elimination. */
int opem_system_matrix(opem_ctx *ctx, double freq, const double *gamma, double *vals);

/* The batched operator of a sweep on caller vectors (parity check of the product SpMV):
 * y_f = A(freqs[f]) x_f for f < nfreq <= 256, A(w) as opem_system_matrix (gamma: complex [nfreq][nports],
 * ports in ascending attribute order), applied as Kt - k0^2 (Mr - j Mi) + boundary terms over one pass
 * of the frequency-independent matrices (the MatMult of KSPSolve, solveComplexLinearSystem.c:100-190,
 * for every frequency at once).  prec 0 FP64, 1 FP32 values and vectors.  x, y complex [nfreq][N]. */
int opem_sweep_apply(opem_ctx *ctx, int prec, int nfreq, const double *freqs, const double *gamma,
                     const double *x, double *y);

/* Complex CSR SpMV on caller-provided host arrays (parity check of the device kernel):
 * y = A x with A n x n (rowptr n+1, colind nnz, vals nnz complex); prec OPEM_PREC_FP64 computes in
 * FP64, OPEM_PREC_MIXED stores values and vectors in FP32 and accumulates in FP64. */
int opem_spmv(opem_ctx *ctx, int prec, int64_t n, const int32_t *rowptr, const int32_t *colind,
              const double *vals, const double *x, double *y);

/* Bits of the flags output of opem_solve / opem_hfield. */
#define OPEM_SOLVE_FP64_FALLBACK 1 /* mixed precision: the FP32 refinement stagnated (true residual down by less */
                                   /* than 10% in two consecutive cycles) and FP64 GMRES finished the solve     */
#define OPEM_SOLVE_SINGULAR 2      /* the sparse-LU factorization met exactly zero pivots (replaced by a tiny    */
                                   /* value): the matrix is singular or nearly so; GMRES still decides           */

/* Options of a caller-matrix Krylov solve (zero fields take the default). */
typedef struct {
  int precision;   /* OPEM_PREC_FP64: FP64 factor and GMRES; OPEM_PREC_MIXED: FP32 factor and inner GMRES cycles   */
                   /*    inside FP64 iterative refinement                                                           */
  int restart;     /* GMRES(m), m <= 400 (0 -> 30)                                                                  */
  int maxit;       /* iteration limit (0 -> 10000, PETSc's default)                                                 */
  double tol;      /* true relative residual ||b - A x|| / ||b|| (0 -> 1e-10)                                       */
  int first_rung;  /* preconditioner rung to start the retry ladder at: 0 sparse LU, 1 point Jacobi, 2 none        */
} opem_solve_opts;

/* solveComplexLinearSystem (src/OpenParEM3D/solveComplexLinearSystem.c:100-209) on a caller's complex CSR matrix
 * (e.g. the AIJ that hypre_ParCSRMatrixToMat builds, csr.c:171-223, after eliminatePEC): b = -A x, b[PortDof] =
 * x[PortDof], the PortDof rows and columns eliminated (MatZeroRowsColumns, diagonal 1), then GMRES preconditioned
 * by a sparse LU of the eliminated matrix (nested dissection of its graph, multifrontal factor on the GPU: the
 * reference's PCCHOLESKY/PCLU), with the reference's retry ladder: when a rung breaks down (non-finite residual)
 * the next one is tried: point Jacobi (the reference's PCBJACOBI rung), then no preconditioner; a rung that runs
 * to maxit ends the ladder (KSP_DIVERGED_ITS).  The pattern must be structurally symmetric, as MFEM assembles it.
 * x: in: the Dirichlet values on port_dof (other entries ignored); out: the solution.  rnorm: ||b - A x||_2 in
 * FP64 (the reference returns KSP's residual norm).  iters: GMRES iterations; converged: 1 when rnorm <= tol ||b||
 * (nonconvergence is a flag, not an error); rung: the ladder rung that ran last; flags: OPEM_SOLVE_* bits (FP64
 * continuation after FP32 stagnation, zero pivots met).  Any of the outputs but x may be NULL. */
int opem_solve(opem_ctx *ctx, int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
               int64_t nport, const int32_t *port_dof, const opem_solve_opts *opts, double *x, double *rnorm,
               int32_t *iters, int32_t *converged, int32_t *rung, int32_t *flags);

/* solveHfield (solveComplexLinearSystem.c:235-315): P h = b for a caller's complex CSR P (the ND mass matrix of
 * build_P, fem3D.cpp:192-207; no boundary conditions) and b (buildHb: b = j Q x/(w mu), solveComplexLinearSystem.c
 * :211-221, which a caller forms with opem_spmv), by the same preconditioned GMRES and retry ladder as opem_solve.
 * h: the solution (n complex). */
int opem_hfield(opem_ctx *ctx, int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
                const double *b, const opem_solve_opts *opts, double *h, double *rnorm, int32_t *iters,
                int32_t *converged, int32_t *rung, int32_t *flags);

/* Mode::calculateSplits (src/OpenParEM3D/port.cpp:3524-3568) for one port of the context's mesh: the modal E_t and
 * H_t sampled at the port's opem_port_points (complex npts*3 each) are projected onto the port's trace space, and
 * with the 3D E and H DOF vectors x and h (complex ndof each): e0 = m_E^H M_p x, e2 = m_E^H M_p m_E, h0, h2 likewise
 * (M_p the exact trace mass matrix, tangentialInnerProduct port.cpp:3463-3489), C+ = (e0/e2 + h0/h2)/2,
 * C- = (e0/e2 - h0/h2)/2.  splits: complex [6] = {C+, C-, e0, e2, h0, h2}. */
int opem_port_project(opem_ctx *ctx, int attr, const double *mode_E, const double *mode_H, const double *x,
                      const double *h, double *splits);

/* The dominant mode of port attr at freq from the 2D port eigenproblem on the port's triangles: OpenParEM2D's
 * ND_p x H1_p generalized eigenproblem (src/OpenParEM2D/fem2D.cpp:2221-2343, eigensolve.c:629-1017: A x = gamma^2 B x
 * with x = (gamma e_t, e_z), the eigenvalue of smallest real part) and its H field (Hsolve.c:454), on the context's
 * mesh, space, materials (of the port faces' owner tets) and boundaries (PEC and surface-impedance conductors are
 * Dirichlet).  e_t is computed in the 3D space's own trace on the port, so the mode drives opem_sweep without a field
 * transfer.  Outputs: gamma = alpha + j beta (complex), E_t and H_t (H_t of the wave leaving the domain) sampled at
 * the port's opem_port_points (complex npts*3 each: the mode_E / mode_H layout of opem_sweep), normalised to a
 * peak |E_t| of 1; with vpath (6 doubles: start and end point of a voltage path in the port) V = int E_t . dl along
 * it (made real and positive), else V = 1; Z = V V* / (2 conj(P)), P = 1/2 int (E_t x H_t*) . n (Z_PV,
 * fem2D.cpp:1819).  The eigenproblem is solved by shift-invert inverse iteration and Rayleigh-quotient steps with
 * the sparse LU on the GPU. */
int opem_port_mode(opem_ctx *ctx, int attr, double freq, const double *vpath, double *gamma, double *mode_E,
                   double *mode_H, double *V, double *Z);

/* ---- Radiation patterns (src/OpenParEM3D/port.cpp:1043-1265, 8077-8160; pattern.cpp) ----------------------- */

/* Boundary::calculateRadiationCurrents (port.cpp:1043-1265) for the radiation boundaries attrs[0..nattr) (nattr 0:
 * every radiation boundary of the context), which must share one wave impedance eta (fem3D.cpp:1904-1919):
 * J = n x H_t and M = -n x E_t (n the outward unit normal) at the centres of the boundary triangles refined until
 * no edge is longer than limit wavelengths (lambda = eta / (mu0 freq), fem3D.cpp:1926-1929), each with its area.
 * x, h: the E and H DOF vectors (complex ndof, opem_sweep's x_out / h_out layout).  Each face whose longest edge
 * is L is split into s*s similar triangles, s = ceil(L / (limit lambda)) (the reference bisects with MFEM's
 * GeneralRefinement; both are midpoint rules of the same surface integrals); the centres follow the face order,
 * and within a face i = 0..s-1, j = 0..s-1-i: the triangle at lattice corner (i,j), then, when i+j <= s-2, the
 * inverted one next to it.  Call with xyz == NULL to get the count in *ncur, then with arrays of that size:
 * xyz (ncur*3), J, M (complex ncur*3), area (ncur).  eta_out (optional): the boundaries' wave impedance. */
int opem_radiation_currents(opem_ctx *ctx, int nattr, const int32_t *attrs, double freq, const double *x,
                            const double *h, double limit, int64_t *ncur, double *xyz, double *J, double *M,
                            double *area, double *eta_out);

/* BoundaryDatabase::calculateFarField (port.cpp:8077-8160; Balanis 1989 pp. 288-289) on the GPU: the far E and H
 * of the currents at the points pts (npts*3) about center: fields = complex npts*4 in the order E_phi, E_theta,
 * H_phi, H_theta of OPEMpoint::set_fields, with theta, phi of each point about center.  beta = 2 pi freq mu0 / eta
 * (fem3D.cpp:1926).  kernel_ms (optional): the GPU time of the summation kernels. */
int opem_far_field(opem_ctx *ctx, int64_t ncur, const double *xyz, const double *J, const double *M,
                   const double *area, const double *center, double beta, double eta, int64_t npts,
                   const double *pts, double *fields, double *kernel_ms);

/* Sphere::create, refine, setMetrics, allocateAreasToPoints (pattern.cpp:618-800), host only: the refined
 * icosahedron about the origin (refine 0..9), its points (the vertices in creation order, then one centre per
 * triangle), point areas (vertices only; the centres carry 0), triangles and spherical triangle areas about
 * center.  Call with pts == NULL for the counts *npts, *ntri. */
int opem_pattern_sphere(double radius, const double *center, int refine, int64_t *npts, double *pts,
                        double *point_area, int64_t *ntri, int32_t *tris, double *tri_area);

/* Circle::create (pattern.cpp:1468-1504), host only: npoints points of the circle of the given latitude about
 * center in the x-y plane, rotated by theta about y and then phi about z (about the origin, as rotate_point). */
int opem_pattern_circle(double radius, const double *center, double theta, double phi, double latitude,
                        int64_t npoints, double *pts);

/* Pattern metrics (pattern.hpp:98, pattern.cpp:858-868, 1144-1176, 2529-2535), host only: point power
 * 0.5 (E_theta H_phi* - E_phi H_theta*), radiated power sum(power * point_area) (complex [1]), per-point gain
 * Re(power / (accepted / total_area)) and directivity Re(power / (radiated / total_area)) (optional, npts each),
 * summary = {max gain dB, max directivity dB, radiation efficiency Re(radiated / accepted)}. */
int opem_pattern_metrics(int64_t npts, const double *fields, const double *point_area, double total_area,
                         const double *accepted_power, double *radiated_power, double *gain, double *directivity,
                         double *summary);

/* BoundaryDatabase::calculateAcceptedPower (port.cpp:7475-7497), host only: |Cm|^2 Pz at the driven S-port minus
 * |Cp|^2 Pz of the driven port's mode under every other driving set.  Cp, Cm: opem_sweep's layout for one
 * frequency (complex nports*nports, [driving set][port]); Pz: the driven port's modal power (complex); sport
 * 0-based.  (The reference evaluates this inside its driving-set loop, before the later sets are solved; here the
 * splits of every set are available.) */
int opem_accepted_power(int nports, const double *Cp, const double *Cm, const double *Pz, int sport,
                        double *accepted);

/* ---- Mesh errors for adaptive refinement (src/OpenParEM3D/fem3D.cpp:1162-1380, OPEM_L2ZZErrorEstimator.cpp) ---- */

/* fem3D::calculateMeshErrors with OPEM_L2ZZErrorEstimator (OPEM_L2ZZErrorEstimator.cpp:18-164) for nsys H fields
 * h (complex nsys*ndof, one per driven port; opem_sweep's h_out layout): per element K and field, the flux
 * (1/mu_r) curl H on K (CurlCurlIntegrator(Inv_mur), fem3D.cpp:1167) is L2-projected onto ND_p by CG on the mass
 * matrix to the relative residual tol (the reference: 1e-12), the error is int_K |s - flux| for the real and the
 * imaginary part (norm p = 1), combined as sqrt(e_re^2 + e_im^2) (fem3D.cpp:1221-1224), and errors (ne) keeps the
 * largest over the fields (fem3D.cpp:1263-1267); errors_sys (optional, ne*nsys) each field's.  The integral uses a
 * collapsed Gauss-Jacobi rule exact to degree 2p+3 (MFEM's rule differs: the integrand |.| is not polynomial). */
int opem_mesh_errors(opem_ctx *ctx, int nsys, const double *h, double tol, double *errors, double *errors_sys);

/* The marking of fem3D::calculateMeshErrors (fem3D.cpp:1300-1325), host only: the count = NE * fraction
 * (truncated, at least 1) elements of largest error, largest first (equal errors in ascending element order),
 * into elements[count] (optional), and the largest error (maxMeshError, optional). */
int opem_mark_elements(int64_t ne, const double *errors, double fraction, int64_t *count, int32_t *elements,
                       double *max_error);

/* Conforming local refinement of the marked tets (ParMesh::GeneralRefinement in fem3D::refineMesh,
 * fem3D.cpp:1384-1447), host only: each marked tet not yet split is bisected at the midpoint of its longest edge,
 * and every tet and boundary triangle holding that edge is split with it, so the mesh stays conforming; children
 * keep their parent's attribute and orientation.  (MFEM bisects along its own refinement edges; the refined meshes
 * differ, both conforming.)  Outputs: vertices (the input ones first), tets, boundary faces and attributes;
 * parent (optional, nt_out): the input tet of each output tet.  Call with xyz_out == NULL for the counts. */
int opem_refine_mesh(int64_t nv, const double *xyz, int64_t nt, const int32_t *tets, const int32_t *tet_attr,
                     int64_t nbf, const int32_t *bfaces, const int32_t *bface_attr, int64_t nmarked,
                     const int32_t *marked, int64_t *nv_out, double *xyz_out, int64_t *nt_out, int32_t *tets_out,
                     int32_t *tet_attr_out, int64_t *nbf_out, int32_t *bfaces_out, int32_t *bface_attr_out,
                     int32_t *parent);

/* Allocation-free variant of opem_spmv (same arguments, same kernel, identical y).  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The caller's rowptr, colind, vals, x and y are staged in a 64 MiB
 * device array that the library holds on each GPU it runs on until the process ends; no hipMalloc, hipFree or
 * allocating library routine is called.  Problems whose staged bytes exceed 64 MiB (FP64: 4 (n+1) + 20 nnz +
 * 32 n bytes, each piece rounded up to 256; FP32 adds 8 nnz + 16 n) are rejected with -2.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (two contexts, or two host threads) race. */
int opem_spmv_noalloc(opem_ctx *ctx, int prec, int64_t n, const int32_t *rowptr, const int32_t *colind,
                      const double *vals, const double *x, double *y);

/* Element matrices of ne independent tets (the assembly kernel alone):
 * A_e = K_e - k0^2 er (1 - j tand) M_e, written as {Re A_e (nd*nd row-major), Im A_e} per element.
 * verts: ne*12 (4 vertices per tet, element-local order as given), er/tand: ne.
 * prec 0 FP64 MFMA, 1 FP32 MFMA, 2 BF16-input MFMA (per-element scaled), FP32 output for 1 and 2.
 * The kernel is run nrep times on device-resident inputs; *avg_seconds is its mean duration
 * (HIP events).  out (ne*2*nd*nd doubles) may be NULL.  Counterpart: the element loop inside MFEM's
 * CurlCurlIntegrator / VectorFEMassIntegrator::AssembleElementMatrix used by fem3D.cpp:115-117. */
int opem_element_matrices(opem_ctx *ctx, int order, int prec, int64_t ne, const double *verts, const double *er,
                          const double *tand, double k0, int nrep, double *out, double *avg_seconds);

/* Times the batched operator of a sweep (the SpMV inside GMRES, PETSc MatMult in KSPSolve,
 * solveComplexLinearSystem.c:100-190) on the context's pattern with device-resident synthetic inputs:
 * S = ngroups*R systems A_s = Kt + c_s M + sum_t d_{s,t} T_t over one pass of the frequency-independent
 * matrices, x system-interleaved; prec 0 FP64, 1 FP32.  Needs a space and, for the boundary terms, the
 * boundaries set.  *avg_seconds = mean kernel duration over nrep launches (HIP events on the context
 * stream); *alg_bytes = algorithmic HBM bytes of one launch:
 *   4 (N+1) + 4 nnz + 2 s_r nnz + 4 (N+1) + 4 nov + s_r nterm nov + R N + 2 S N s_v
 * (s_r = 8 FP64 / 4 FP32 real, s_v = 16 / 8 complex; nov, nterm = boundary-term entries and terms). */
int opem_spmv_bench(opem_ctx *ctx, int prec, int ngroups, int R, int nrep, double *avg_seconds, double *alg_bytes);

/* Times the single-matrix complex CSR SpMV y = A x (MatMult of KSPSolve on the explicit complex AIJ matrix,
 * solveComplexLinearSystem.c:100-190, stored as hypre_ParCSRMatrixToMat builds it, csr.c:171-223) on the
 * context's pattern: explicit complex values, one vector, device-resident synthetic inputs.  Each launch reads
 * its own copy of the column indices and values (copies in rotation, at least 2 GiB in all) so no launch is
 * served from an earlier one's cache lines.  prec 0 FP64, 1 FP32.
 * row_group: 0 = the largest G in {3, 2, 1} such that every G consecutive rows share one column list (ND2: 2,
 * ND3: 3, ND1: 1), 1 = one row at a time, 2 or 3 = that G (misuse if the pattern does not admit it); the kernel
 * reads each column index and gathers each x value once per group.  lanes_per_row 0 = default (16 for groups,
 * else the context's choice from the mean row length), else 4/8/16/32/64 (8/16/32 for groups).
 * *avg_seconds = mean kernel duration over nrep launches (HIP events on the context stream); *alg_bytes = the
 * bytes the kernel must move, nnz s_v + 4 nnz/G + 4 (N+1) + 2 N s_v (s_v = 16 FP64 / 8 FP32 complex); with G = 1
 * this is the plain CSR count nnz (s_v + 4) + 4 (N+1) + 2 N s_v. */
int opem_csr_spmv_bench(opem_ctx *ctx, int prec, int row_group, int lanes_per_row, int nrep, double *avg_seconds,
                        double *alg_bytes);

/* Times the Arnoldi orthogonalisation kernels of the batched GMRES (classical Gram-Schmidt inside
 * KSPGMRES, the VecMDot / VecMAXPY pair of solveComplexLinearSystem.c's KSPSolve) for S systems, a basis
 * of k vectors and ns new vectors per pass (1 = one-step Arnoldi, 2..4 = s-step): avg_seconds[0] =
 * projection C = V^H [z_1..z_ns], avg_seconds[1] = update Z -= V C (+ norms / Gram matrix);
 * alg_bytes[0] = S (k+ns) N s_v + S nblk k ns 16, alg_bytes[1] = S (k+2ns) N s_v + S k ns 16 + S nblk' g 8
 * (nblk = 4 ceil(N / 2048) projection partials, nblk' = ceil(N / 1024) update partials of g doubles). */
int opem_krylov_bench(opem_ctx *ctx, int prec, int S, int k, int ns, int nrep, double *avg_seconds,
                      double *alg_bytes);

typedef struct {
  int nfreq;
  const double *freqs;          /* [nfreq] Hz                                                     */
  int nports;                   /* S-ports (one mode per port)                                    */
  const int32_t *port_attr;     /* [nports] boundary attributes, S-port order                     */
  const double *mode_E;         /* complex [nfreq][sum_k npts_k*3] modal E_t at port points       */
  const double *mode_H;         /* complex, same layout: modal H_t of the wave leaving the domain */
  const double *gamma;          /* complex [nfreq][nports] propagation constant alpha + j beta    */
  const double *V;              /* complex [nfreq][nports] modal voltage                          */
  const double *Z;              /* complex [nfreq][nports] modal impedance                        */
  int precision;                /* OPEM_PREC_FP64 or OPEM_PREC_MIXED                              */
  int batch;                    /* frequencies per batched launch; 1 = one system group at a time, */
                                /*    0 = all; lowered to at most 256 systems (batch x nports) and */
                                /*    to what three quarters of the device memory holds           */
  int restart;                  /* GMRES(m), m <= 400 (0 -> 150)                                  */
  int maxit;                    /* Krylov iteration limit per system                              */
  double tol;                   /* true relative residual ||b - A x|| / ||b|| target              */
  int reciprocal;               /* 1: S <- (S + S^T)/2 (results.cpp:401-419)                      */
  double h_tol;                 /* H-projection CG relative tolerance (0 -> 1e-13)                */
  int precond;                  /* 0: vertex-patch additive Schwarz (element blocks if a patch has */
                                /*    more than 100 DOFs), 1: element-block Jacobi, 2: sparse LU   */
                                /*    (nested-dissection multifrontal factor of each frequency's  */
                                /*    matrix, FP32 fronts in mixed precision: the PCCHOLESKY/PCLU */
                                /*    of solveComplexLinearSystem.c:133-168)                      */
  int coarse_size;              /* aggregates of the gradient coarse level (0 = one-level, <= 128) */
  int orth_passes;              /* classical Gram-Schmidt passes per Arnoldi step: 2 (default) or 1 */
  int precond_share;            /* adjacent frequencies of a batch sharing one set of patch inverses */
                                /*    or one LU factor, built from their mean matrix (0 -> 1; x nports */
                                /*    <= 32 for patches in mixed precision, <= 8 in FP64, <= 64 for LU; */
                                /*    LU uses the largest divisor of the batch's frequency count)   */
  int poly_degree;              /* d > 0: GMRES-polynomial preconditioning, the Krylov space built with  */
                                /*    B p(B) (B = A M, deg p = d - 1, roots from a d-step Arnoldi run per */
                                /*    system); each iteration applies B d times (0 = off, <= 32)       */
  int arnoldi_step;             /* s = 2..4: s Krylov vectors per Gram-Schmidt pass (z_1 = B v,        */
                                /*    z_{c+1} = B z_c, block projection, Cholesky QR of the block);     */
                                /*    0 or 1: one; with poly_degree > 0 one is used                    */
  int warm_start;               /* L = 1..4 (with batch >= nfreq >= 3): bisection over the frequency    */
                                /*    index, every 2^L-th frequency first, then the midpoints, each     */
                                /*    started from the linear interpolation in frequency of its solved  */
                                /*    neighbours (free rows); 0: off                                   */
  double mem_fraction;          /* share of the device memory the batches of this sweep may fill by the */
                                /*    planner's estimate (0 -> 0.75); contexts that sweep on one GPU at */
                                /*    the same time divide it between them                             */
} opem_sweep_in;

typedef struct {
  double *S;                    /* complex [nfreq][nports][nports], S[i][j] row-major             */
  double *S_raw;                /* optional: before forced reciprocity                            */
  double *Cp, *Cm;              /* optional complex [nfreq][nports drive][nports]                 */
  int32_t *iters;               /* optional [nfreq][nports]                                       */
  double *true_res;             /* optional [nfreq][nports] final ||b-Ax||/||b|| in FP64          */
  int32_t *converged;           /* optional [nfreq][nports]                                       */
  double *x_out;                /* optional complex [nfreq][nports][ndof]  E DOFs                 */
  double *h_out;                /* optional complex [nfreq][nports][ndof]  H DOFs                 */
  double seconds_total;         /* wall time of the sweep (device-synchronised)                   */
  double seconds_setup;         /* coefficient combine + preconditioner build                     */
  double seconds_solve;         /* E solves                                                       */
  double seconds_post;          /* H projection + port projections + S                            */
  int32_t *fallback;            /* optional [nfreq][nports]: 1 when the FP32 inner solves of a    */
                                /*    mixed-precision system stagnated (its true residual fell by */
                                /*    less than 10% in two consecutive refinement cycles) and FP64 */
                                /*    GMRES finished it from its iterate (iters counts both)      */
  int32_t *singular;            /* optional [nfreq][nports]: 1 when the sparse-LU factorization  */
                                /*    of the system's batch met exactly zero pivots (replaced by a */
                                /*    tiny value; the factor is then only a preconditioner of a   */
                                /*    singular or near-singular matrix and GMRES decides)         */
} opem_sweep_out;

/* The per-frequency hot path for a list of frequencies and all port excitations:
 * coefficients -> A(omega) -> PEC + driven-port elimination -> GMRES (element-block Jacobi)
 * -> H projection -> port projections (C+, C-) -> S.  Mirrors the driving-set loop of
 * OpenParEM3D.cpp:497-566 around fem3D::solve (fem3D.cpp:712-799) and
 * BoundaryDatabase::calculateS (port.cpp:7499-7606). */
int opem_sweep(opem_ctx *ctx, const opem_sweep_in *in, opem_sweep_out *out);

/* Kernel timings of the sparse-LU preconditioner (precond 2) on this context, accumulated since the last reset:
 * the factorization's trailing updates, panels and the apply's matrix products, each timed with HIP events on the
 * context stream while the environment variable OPEM_LU_TIMING is set (counts only otherwise).  out[15]:
 *   [0] trailing-update seconds on the 64-tile kernel  [1] their real flops (8 per complex multiply-add)
 *   [2] launches  [3] panel seconds  [4] launches  [5] apply matrix-product seconds  [6] matrix bytes they read
 *   [7] launches  [8] factorization wall seconds  [9] matrices factored  [10] matrix-applies  [11] 1 when event
 *   timing is on  [12] deep FP32 trailing-update seconds on the 128-tile kernel  [13] their real flops  [14] launches.
 * reset != 0 clears the counters after reading.  Counterpart: the PCFactor/MUMPS timing a PETSc -log_view
 * reports for the PCCHOLESKY/PCLU of solveComplexLinearSystem.c:133-168. */
int opem_lu_stats(opem_ctx *ctx, int reset, double *out);

/* S from splits (BoundaryDatabase::calculateS, port.cpp:7499-7606, with Zsolve.c LUP inverse):
 * Cp/Cm complex [n drive][n], V/Z complex [n]; S complex [n][n]. Host-only, no context. */
int opem_calculate_s(int n, const double *Cp, const double *Cm, const double *V, const double *Z,
                     int reciprocal, double *S);

/* ---- S-parameter post-processing (host only, no context; src/OpenParEM3D/results.cpp, process3D.cpp) ----------
 * S matrices are complex [nports][nports] row-major per frequency, the layout of opem_sweep_out.S (the reference's
 * Result::S Mat).  These calls return 0, -2 for misuse, 1 for a failure (singular matrix, unwritable file);
 * opem_results_error() gives the message of the calling thread's last failed call. */
const char *opem_results_error(void);

/* Result::S2Z (results.cpp:455-511): Z = k (I + S)(I - S)^-1 k, k = diag(sqrt(Zo)); Zo complex [nports]. */
int opem_s2z(int nports, const double *S, const double *Zo, double *Z);
/* Result::Z2S (results.cpp:513-560): S = (Z k^-1 + k)^-1 (Z k^-1 - k). */
int opem_z2s(int nports, const double *Z, const double *Zo, double *S);
/* Result::renormalize (results.cpp:562-578): S2Z with the ports' Zo, then Z2S with every port at Zref. */
int opem_renormalize(int nports, const double *S, const double *Zo, double zref_re, double zref_im, double *S_out);

/* What the S-parameter files say besides the numbers (the reference takes them from its project data). */
typedef struct {
  const char *version;          /* "major.minor.patch" for the first line; NULL or "" omits the line            */
  const char *format;           /* "RI", "MA" or "DB" (touchstone_format)                                      */
  const char *unit;             /* "Hz", "kHz", "MHz" or "GHz" (touchstone_frequency_unit)                     */
  double reference_impedance;   /* ohm after renormalisation; 0: not renormalized                              */
  const char *const *nets;      /* [nports] net names; NULL (or a NULL entry): net1, net2, ...                 */
} opem_sparam_file;

/* ResultDatabase::saveCSV (results.cpp:940-1046, rows Result::saveCSV :208-232): header lines, then one line per
 * frequency with the entries (i, j) in the order i outer, j inner; freqs in Hz. */
int opem_write_csv(const char *path, const opem_sparam_file *meta, int nports, int64_t nfreq, const double *freqs,
                   const double *S);
/* ResultDatabase::saveTouchstone, version 1.1 (results.cpp:1048-1190): rejects reference_impedance == 0, as the
 * reference skips the file for data that is not renormalized.  Mixed-mode (Touchstone 2.0) output is not built. */
int opem_write_touchstone(const char *path, const opem_sparam_file *meta, int nports, int64_t nfreq,
                          const double *freqs, const double *S);
/* ResultDatabase::save_as_test (results.cpp:276-341, 1540-1580): appends prototype regression cases for the
 * results (magS(dB) and argS(deg), "equal" to 1e-12 or "lessthan" below -150 dB / 0.1 deg); *casenumber counts on. */
int opem_save_as_test(const char *path, const char *casename, int nports, int64_t nfreq, const double *freqs,
                      const double *S, int *casenumber);
/* process3D (process3D.cpp:481-559): evaluates the S-parameter test cases in cases_path against a results CSV in
 * DB format (as opem_write_csv writes it) and writes the sorted report to report_path.  The counts come back in
 * the outputs: all cases passed when *npassed == *ncases (process3D's exit status 0). */
int opem_process_tests(const char *cases_path, const char *results_csv, const char *report_path, int *ncases,
                       int *nevaluated, int *npassed);

#ifdef __cplusplus
}
#endif
#endif /* OPEM_H */
