/*
 * acp.h -- C ABI of the MI355X D3Q19 lattice-Boltzmann engine (libacpulse.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  This is
 * the boundary a caller of the reference solver (aleksandrShukman/acPulse_vcLBM)
 * binds instead of its Control / Lattice classes; each entry point names the
 * reference interface it replaces.  Populations are exchanged as float64 arrays
 * in the layout f[q][k][j][i] (q in the reference's direction order, i fastest),
 * i.e. index q*N + i + nx*(j + ny*k) with N = nx*ny*nz.
 *
 * Threading: one host thread per case handle at a time; different handles are
 * independent.  All device work of a case is issued on the stream given to
 * acp_case_create (or a stream the library creates when NULL is passed).
 */
#ifndef ACP_H
#define ACP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference exits the process on configuration errors
 * (Control.cpp:147, 164; Node.cpp:206, 436); this library returns them. */
enum {
    ACP_OK = 0,
    ACP_E_ARG = 1,         /* invalid argument or unsupported parameter combination */
    ACP_E_CONFIG = 2,      /* case file unreadable, key missing, bad value */
    ACP_E_OOM = 3,         /* device memory allocation failed */
    ACP_E_HIP = 4,         /* a HIP runtime call failed */
    ACP_E_STATE = 5,       /* call out of order (e.g. step before init) */
    ACP_E_UNSUPPORTED = 6, /* feature not available in this build (e.g. acp_case_create with MaxLevel > 0) */
    ACP_E_IO = 7           /* output file could not be written */
};

/* Collision-Model (BoundaryConditions.txt:31-44; Lattice.cpp:2158-2183) */
enum { ACP_OP_SRT = 0, ACP_OP_MRT_GS = 1, ACP_OP_MRT_RM = 2, ACP_OP_RR = 3, ACP_OP_HRR = 4 };
/* Storage / arithmetic tier.  F64 reproduces the reference arithmetic; the shifted
 * tiers store f_i - w_i rho0 (scaled) and collide in float32, in lattice units. */
enum { ACP_PREC_F64 = 0, ACP_PREC_F32_SHIFTED = 1, ACP_PREC_F16_SHIFTED = 2 };
/* Domain boundary: the reference applies LODI on all six faces (Lattice.cpp:1582-1624).
 * PERIODIC wraps all three axes.  The other kinds wrap y and z and give each x
 * face its own rule:
 *   LODI_X       LODI on both x faces, as the reference treats its x faces
 *                (Lattice.cpp:1586-1593, Node.cpp:229-504; no edges or
 *                corners, so every face node has its four tangent neighbours,
 *                wrapped in y and z); at least 4 nodes in x.  A duct with
 *                non-reflecting ends: the coarse level of the refinement
 *                measurement case.
 *   OPEN_X       both x faces open: their nodes collide like interior nodes and
 *                links leaving the grid are dropped, so something else must set
 *                them (the fine level of a refined case, acp_refined_*).
 *   OPEN_LODI_X  face x = 0 open, face x = nx - 1 LODI: the fine level of a
 *                refined case whose region reaches the high-x face of a
 *                LODI_X coarse level (acp_refined_create makes it).
 * HRR's strain stencil, and the cubic-Mach terms' (RR and HRR), take the
 * missing neighbour beyond an open face by quadratic extrapolation
 * (u(-1) = 3 u(0) - 3 u(1) + u(2), the second-order one-sided difference). */
enum { ACP_BC_LODI = 0, ACP_BC_PERIODIC = 1, ACP_BC_OPEN_X = 2, ACP_BC_LODI_X = 3, ACP_BC_OPEN_LODI_X = 4 };
/* MRT operators: the reference computes the relaxed populations and discards them
 * (Node.cpp:2603-2604, 2833-2834), i.e. MRT steps are pure streaming + LODI. */
enum { ACP_MRT_AS_REFERENCE = 0, ACP_MRT_RELAXED = 1 };

/* Case description: the fields of Control (Control.h:77-127) plus the options
 * this engine adds.  Lengths in metres, rates in SI as in the case file. */
typedef struct acp_case_desc {
    int32_t nx, ny, nz;         /* nodes (Control.cpp:43-45) */
    int32_t op;                 /* ACP_OP_* */
    int32_t order;              /* Order-Of-Equilibrium 2|3|4 */
    int32_t cubic;              /* Cubic-Mach-Correction yes=1 */
    int32_t transverse;         /* Transverse yes=1 */
    int32_t nonreflecting;      /* Non-Reflecting yes=1 (selects arrays in VTK output only) */
    int32_t boundary;           /* ACP_BC_* */
    int32_t mrt_mode;           /* ACP_MRT_* */
    int32_t precision;          /* ACP_PREC_* */
    int32_t max_level;          /* MaxLevel; 0 only */
    int32_t timesteps;          /* Max-Time-Step */
    int32_t write_interval;     /* WriteInterval */
    int32_t threads;            /* Threads (kept for output parity; unused) */
    double rho0;                /* Density */
    double nu;                  /* KinematicViscosity */
    double cs;                  /* SpeedOfSound */
    double spacing;             /* Spacing */
    double mach;                /* MACH-Number */
    double k1;                  /* K1-Relaxation */
    double k2;                  /* K2-Relaxation (parsed, unused as in the reference) */
    double sigma;               /* Hybrid-Parameter */
    double domain[3];           /* Domain-Size-X/Y/Z */
    double f16_scale;           /* storage scale of the FP16 tier (per lattice) */
    char save_as[1024];         /* Save-as (rest of line) */
    char collision_model[32];   /* Collision-Model as written */
} acp_case_desc;

typedef struct acp_case acp_case;

/* Case handle facts for admission and reporting. */
typedef struct acp_case_info {
    int64_t nodes;
    int64_t device_bytes;       /* device memory held by the case */
    int64_t bytes_per_node_step;/* algorithmic HBM bytes of the bulk kernel per node and step */
    int64_t lodi_nodes;         /* boundary nodes handled by the LODI kernel */
    int32_t lodi_launches;      /* LODI kernel launches per step */
    int32_t steps_done;
    double omega;               /* collision frequency (Control.cpp:54) */
    double dt;                  /* time step (Control.cpp:53) */
} acp_case_info;

/* Last error text of the calling thread ("" when none). */
const char* acp_last_error(void);

/* Control::Control(string) (Control.cpp:8-132): parse a BoundaryConditions.txt
 * with the reference's first-match keyword semantics and derive node counts.
 * Also reads the optional keys this engine adds (Precision, Boundary, MRT-Mode,
 * FP16-Scale); unknown keys are ignored as in the reference. */
int acp_control_read(const char* path, acp_case_desc* out);
/* Control::readValue / readString / readFileName / readVector
 * (Control.cpp:135-273) for any keyword of a case file: the first line whose
 * first token is `keyword`; the number, the next token, the rest of the line,
 * or three numbers.  A missing file or keyword is ACP_E_CONFIG with the
 * reference's message (the reference exits); a string longer than cap - 1
 * bytes is ACP_E_ARG. */
int acp_control_read_value(const char* path, const char* keyword, double* out);
int acp_control_read_string(const char* path, const char* keyword, char* out, int64_t cap);
int acp_control_read_filename(const char* path, const char* keyword, char* out, int64_t cap);
int acp_control_read_vector(const char* path, const char* keyword, double* out3);

/* Defaults of the reference's BoundaryConditions.txt for every field. */
void acp_case_desc_default(acp_case_desc* out);

/* Lattice::Lattice(Control&) (Lattice.h:23-57): allocate device state for one
 * uniform-level case.  `stream` is a hipStream_t (NULL: the library creates one). */
int acp_case_create(const acp_case_desc* desc, void* stream, acp_case** out);
void acp_case_destroy(acp_case* c);
/* The same inside a caller-owned device region (256-byte aligned, at least
 * acp_case_footprint bytes): the case allocates nothing and frees nothing, so
 * creating and destroying it never waits on other streams.  No counterpart in
 * the reference; for servers that keep many cases resident. */
int acp_case_create_in(const acp_case_desc* desc, void* stream, void* region, int64_t region_bytes, acp_case** out);
int acp_case_info_get(const acp_case* c, acp_case_info* out);

/* Lattice::initialization(Control&) (Lattice.cpp:1397-1493): the reference's
 * hard-coded Gaussian pulse, feq_N(rho, 0) in both arrays. */
int acp_case_init_reference(acp_case* c);
/* Same initialization with a parameterised pulse in node units:
 * rho = rho0 (1 + amp exp(-ln2 ((i-x0)^2 + (j-y0)^2) / b^2)), u = 0. */
int acp_case_init_pulse(acp_case* c, double amp, double x0, double y0, double b);
/* Equilibrium state feq_N(rho, 0) from a host density field rho[N] (kg/m^3,
 * node order i + nx*(j + ny*k)), in both arrays; no counterpart in the
 * reference, whose initial field is hard-coded. */
int acp_case_init_density(acp_case* c, const double* rho);
/* Random fluctuation field on the device: rho = rho0 (1 + rho_amp n0),
 * u_a = u_amp c_s sqrt(3) n_a with n ~ N(0,1) from Philox-4x32-10 (key = seed,
 * counter = node index) and Box-Muller; populations feq_N(rho, u). */
int acp_case_init_random(acp_case* c, uint64_t seed, double rho_amp, double u_amp);
/* Replace the populations of both arrays (host float64 [19][N]). */
int acp_case_set_populations(acp_case* c, const double* f);

/* Lattice::solve(bool&) (Lattice.cpp:2158-2183) n times.  Asynchronous on the
 * case stream; errors of earlier work surface on the next synchronising call. */
int acp_case_step(acp_case* c, int32_t n);
/* Block until the case stream is idle. */
int acp_case_sync(acp_case* c);

/* No counterpart in the reference (the reference launches nothing): run every
 * full chunk of `chunk` steps inside acp_case_step as one replay of a captured
 * HIP graph instead of one launch sequence per step.  `chunk` is even (the
 * arrays end where they started) or 0 to switch graphs off.  Results are
 * identical to plain stepping; probes keep recording every step. */
int acp_case_enable_graphs(acp_case* c, int32_t chunk);

/* No counterpart in the reference: the addressing and streaming form of the
 * bulk (collide + push) kernels.  ACP_BULK_AUTO (the default) follows the
 * arrays' size: 32-bit offsets from a uniform base while the per-node part of
 * every offset stays below 4 GiB, block-relative (float64 bulk kernels) or
 * 64-bit offsets beyond; non-temporal loads and stores once the two population
 * arrays exceed the 256 MB last-level cache.  The other values force one form
 * on any grid, so that tests can compare the forms only large grids take (the
 * 512^3 float64 grid runs ACP_BULK_RELATIVE_NT, its wrapped planes
 * ACP_BULK_WIDE_NT) with the oracle on small grids; results are identical in
 * every form.  A 32-bit form on arrays that need 64-bit offsets is ACP_E_ARG.
 * ACP_BULK_RELATIVE_NT forces the float64 kernels' block-relative form: 32-bit
 * offsets from a base a few blocks before each workgroup's own block, which
 * ACP_BULK_AUTO takes on arrays too large for offsets from the array start
 * (the 512^3 float64 grid); rows whose z neighbours wrap to the far end of a
 * periodic grid keep 64-bit offsets.  Grids of a single block, and the
 * shifted tiers, run the 64-bit form under it.
 * ACP_BULK_MFMA is the matrix-core trial of the relaxed MRT-RM collision:
 * f* = f + C (feq - f) with the case's 19 x 19 collision matrix C = M^-1 S M
 * on v_mfma_f64_16x16x4f64 (float64, periodic grids with nx a multiple of
 * 256, operator MRT-RM with ACP_MRT_RELAXED; otherwise ACP_E_UNSUPPORTED).  It
 * equals the default form in exact arithmetic, not to the bit (the matrix
 * product rounds in its own order). */
enum {
    ACP_BULK_AUTO = 0,
    ACP_BULK_WIDE_NT = 1,
    ACP_BULK_NARROW_NT = 2,
    ACP_BULK_NARROW_PLAIN = 3,
    ACP_BULK_MFMA = 4,
    ACP_BULK_RELATIVE_NT = 5
};
int acp_case_set_bulk_variant(acp_case* c, int32_t variant);

/* No counterpart in the reference: device bytes a case of `desc` holds, with
 * `nprobe` probes of `probe_capacity` steps, for admission by HBM footprint.
 * Every buffer counts in multiples of 256 bytes.  Validates `desc` as
 * acp_case_create does; allocates nothing. */
int acp_case_footprint(const acp_case_desc* desc, int32_t nprobe, int32_t probe_capacity, int64_t* bytes);

/* No counterpart in the reference: order-independent 64-bit fingerprint of the
 * current population array in its storage format (bit patterns mixed with the
 * element index and summed modulo 2^64).  Equal states give equal values. */
int acp_case_checksum(acp_case* c, uint64_t* out);

/* Lattice::getDistribution(bool&, ...) (Lattice.h:119) for the whole current
 * array: host float64 [19][N]. */
int acp_case_get_populations(acp_case* c, double* f);
/* The same for the other array of the ping-pong pair (the one the last step
 * read; after initialisation or acp_case_set_populations both hold the same
 * state): Lattice::getDistribution with the flag of the previous step.  Its
 * interior nodes hold their state before the last step's collision; the
 * reference collides in place (Node.cpp:511-557), so its array holds their
 * post-collision values.  Boundary nodes, which do not collide, agree. */
int acp_case_get_previous_populations(acp_case* c, double* f);
/* Density and velocity of every node from the current array (Node.cpp:11-30,
 * 85-118, without cubic-Mach terms): rho [N], u [3][N], float64. */
int acp_case_get_macros(acp_case* c, double* rho, double* u);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): acp_case_get_macros with the same results, staged in a
 * fixed device buffer of the library instead of one allocated per call.  The
 * buffer holds 4 194 304 nodes (128 MiB, held by each GPU the library runs on
 * until the process ends); a larger case is ACP_E_ARG.  Calls that overlap in
 * time on one GPU share that buffer and race: serialise them. */
int acp_case_get_macros_noalloc(acp_case* c, double* rho, double* u);
/* Node::getCubicMachCorrection (Node.h:52) for every node: the cubic-Mach
 * terms psi [19][N] the last step computed (zeros for a case without them). */
int acp_case_get_psi(acp_case* c, double* psi);
/* Total mass sum over all nodes of sum_q f_q, reduced in float64 on the device. */
int acp_case_mass(acp_case* c, double* mass);
/* Probes: nodes whose density (float64) is recorded on the device after every
 * step; read back the recorded series [steps][nprobe] and clear it. */
int acp_case_set_probes(acp_case* c, const int64_t* nodes, int32_t nprobe, int32_t capacity);
int acp_case_read_probes(acp_case* c, double* out, int32_t* nsteps);

/* Lattice::writeResultsVTK(string, bool&, bool&) (Lattice.cpp:2212-2243, 4094-4349)
 * for one level: writes <stem>.vtk with the reference's byte layout from the
 * current state (and, for Non-Reflecting yes, the previous array of boundary
 * nodes with i > 0, as the reference prints them). */
int acp_case_write_vtk(acp_case* c, const char* stem);
/* The same writer on host arrays (float64 [19][N] current and previous arrays,
 * cubic-Mach psi or NULL), e.g. for states computed elsewhere. */
int acp_write_vtk_host(const acp_case_desc* desc, const double* fcur, const double* fprev, const double* psi,
                       const char* stem);

/* Introspection, no counterpart in the reference: the boundary nodes of an
 * nx*ny*nz LODI case by dependency phase (phases derived from the reference's
 * pass order, Lattice.cpp:1582-1624; within a phase by ascending node index)
 * and the phase of each.  Returns the node count; writes at most `cap` entries. */
int64_t acp_lodi_schedule(int32_t nx, int32_t ny, int32_t nz, int64_t* nodes, int32_t* phases, int64_t cap);
/* The same nodes as the LODI kernels receive them, one record each: the node,
 * its normal and second-normal neighbours (whose B entries it reads), four
 * tangent neighbours (-1: none), inward normal direction and corner flag as
 * Lattice::setConnectivity assigns them (Lattice.cpp:500-1392).
 * `launch_start` receives the first record of every launch (nlaunch + 1
 * entries).  Returns the record count; writes at most `cap` records and
 * `launch_cap` launch starts; *nlaunch receives the launch count. */
typedef struct acp_lodi_record {
    int64_t node, n1, n2, t[4];
    int32_t dir, corner;
} acp_lodi_record;
int64_t acp_lodi_launches(int32_t nx, int32_t ny, int32_t nz, acp_lodi_record* records, int64_t cap,
                          int64_t* launch_start, int32_t launch_cap, int32_t* nlaunch);

/* ------------------------------------------------------------------ two-level refinement
 * Lattice::solve(bool&, bool&) with MaxLevel 1 (Lattice.cpp:2185-2209, nested
 * stepping Lattice.cpp:4354-15051) for the coupling the reference's case file
 * selects (Coupling-Approach STD, Overlap 1, Coarse-to-Fine fAll,
 * Fine-to-Coarse cAll, F2C-Filter NOFI, Scaling-Style SRTSTYLE;
 * BoundaryConditions.txt:52-106), whose arithmetic (GridCoupling.cpp) the
 * reference does not contain; restated from the published scheme.  This is
 * the engine's x band (the reference's own fine shell around a coarse core is
 * acp_refined_create_shell): the coarse
 * level is an ACP_BC_PERIODIC or ACP_BC_LODI_X grid and the fine level covers
 * coarse x in [x0, x1] over every y and z at half the spacing and half the
 * time step, with ratio-2 vertex-centred coupling on its x faces:
 *   periodic coarse level: 1 <= x0, x0 + 2 <= x1 <= nx, x1 - x0 < nx - 1
 *     (x1 = nx is the wrapped column 0); both fine faces are coupled;
 *   LODI_X coarse level: 1 <= x0, x0 + 2 <= x1 <= nx - 1.  With x1 < nx - 1
 *     both fine faces are coupled; with x1 = nx - 1 the fine region reaches
 *     the high-x face, the fine level's face there is a LODI face
 *     (ACP_BC_OPEN_LODI_X) and only the face at x0 is coupled.
 * Every coarse node strictly between x0 and x1 is restricted from the fine
 * level.  A fine LODI face uses the level-0 domain length and its own spacing
 * and time step, as the reference's LODIpressureOutlet does on every level
 * (Node.cpp:237-311).  Every operator, every storage tier (both levels use
 * the coarse description's tier); the cubic-Mach correction with RR and HRR
 * in float64.  HRR hands strain rates across the interface as the
 * reference's partner functions compute them (Node.cpp:1892-2490): the fine
 * face nodes collide with their coarse partners' (at t and t + 1/2), the
 * coarse grid-coupling columns with their fine partners'
 * (acp_refined_set_coupling for the direct coupling's form).  With the
 * cubic-Mach correction the same functions hand psi from the same stencils
 * (Node.cpp:2069), and those interface nodes of RR and HRR collide with the
 * handed psi in their moments, Pi1 and result; every other node keeps the
 * psi of its own collision. */
typedef struct acp_refined acp_refined;
int acp_refined_create(const acp_case_desc* coarse, int32_t x0, int32_t x1, void* stream, acp_refined** out);
/* The reference's own two-level geometry (MaxLevel 1; createLattice,
 * Lattice.cpp:6-213, with RefinementLayerYZ = layer_yz, RefinementLayerX =
 * layer_x, Overlap 1): a fine shell around a coarse core.  The coarse level
 * `coarse` has LODI on every face (ACP_BC_LODI); the fine level covers the
 * domain at half the spacing (2 nx - 1, 2 ny - 1, 2 nz - 1 nodes, LODI faces
 * relaxing over the level-0 domain length) outside the hole box
 * [2 layer_x, 2 (nx - 1 - layer_yz)] x [2 layer_yz, 2 (ny - 1 - layer_yz)] x
 * [2 layer_yz, 2 (nz - 1 - layer_yz)], whose surface holds the fine
 * grid-coupling nodes; the coarse grid-coupling nodes are the surface of the
 * core box one coarse cell outside it.  The standard coupling with its link
 * sets (acp_refined_set_coupling; the links a fine grid-coupling node misses
 * are those arriving from inside the hole, setConnectivity's empty
 * neighbours), the filters and Scaling-Styles of acp_refined_*, HRR's strain
 * rates and the cubic-Mach psi handed by the partner functions at faces,
 * edges and corners of the interface.  The fine level does not step the
 * nodes strictly inside the hole box (no part of the reference's level 1;
 * their populations are left as they are and nothing reads them).  The
 * coarse level steps the whole domain: its nodes outside the core box stand
 * for nodes the reference does not step and are overwritten where the
 * coupling reads them (every coarse node outside the core box off the domain
 * faces takes its coincident fine node).  Float64; layers of at least 2 that leave
 * a core of at least 3 nodes per direction; no placement in a region
 * (acp_refined_create_in is for the x band). */
int acp_refined_create_shell(const acp_case_desc* coarse, int32_t layer_yz, int32_t layer_x, void* stream,
                             acp_refined** out);
void acp_refined_destroy(acp_refined* r);
/* The level's case (0 coarse, 1 fine) for probes, read-back and output; step
 * it only through acp_refined_step.  Probes record each level's state after
 * its own step, before the coupling writes into it. */
int acp_refined_level(acp_refined* r, int32_t level, acp_case** out);
/* acp_case_init_pulse on both levels with the same field (coarse node units). */
int acp_refined_init_pulse(acp_refined* r, double amp, double x0, double y0, double b);
/* Lattice::initialization (Lattice.cpp:1397-1493) on both levels: the
 * reference's hard-coded pulse evaluated at each level's own node
 * coordinates (fine node i of the region lies at fine index 2 x0 + i of the
 * domain, spacing / 2), feq_N(rho, 0) in both arrays. */
int acp_refined_init_reference(acp_refined* r);
/* n coarse steps (2n fine steps). */
int acp_refined_step(acp_refined* r, int32_t n);
/* Lattice::writeResultsVTK of a two-level lattice (Lattice.cpp:2212-4092,
 * multi-level branch): <stem>.vtk with the fine level's nodes first, then the
 * coarse nodes outside the fine region's interior (columns 0..x0 and
 * x1..nx-1), VOXEL cells of each level, POINT_DATA Omega, NodeID (0 fluid, 1
 * LODI boundary, 2 grid-coupling), Pressure, Density and Velocity in that
 * point order, numbers as the single-level writer prints them.  A shell case
 * (acp_refined_create_shell) writes the reference's own two-level layout:
 * the level-1 bottom, top, left and right slabs and inlet and outlet blocks
 * of the shell, then the level-0 core, with the reference's cell indexing
 * (the core's cells keep the level-1 plane size as z stride, as the
 * reference's writer does).  With the cubic-Mach terms the density and
 * velocity include psi, as calcDensity does. */
int acp_refined_write_vtk(acp_refined* r, const char* stem);
/* The same from host arrays (the coarse level's description and current /
 * other arrays, [19][N] each; the other arrays only under Non-Reflecting). */
int acp_write_vtk2_host(const acp_case_desc* coarse, const double* ccur, const double* cprev, int32_t x0, int32_t x1,
                        const double* fcur, const double* fprev, const char* stem);
/* The shell geometry's two-level output from host arrays (the psi arrays may
 * be NULL; no counterpart in the reference beyond its writer). */
int acp_write_vtk_shell_host(const acp_case_desc* coarse, int32_t layer_yz, int32_t layer_x, const double* ccur,
                             const double* cprev, const double* cpsi, const double* fcur, const double* fprev,
                             const double* fpsi, const char* stem);
/* The fine level's description of a refined case (acp_refined_create's checks). */
int acp_refined_fine_desc(const acp_case_desc* coarse, int32_t x0, int32_t x1, acp_case_desc* fine);
/* Shifted tiers: run the coupling arithmetic (time interpolation, rescaling,
 * restriction) in float64 on SI populations (ACP_PREC_F64, the default) or in
 * float32 on lattice-unit deviations (ACP_PREC_F32_SHIFTED). */
int acp_refined_set_coupling_precision(acp_refined* r, int32_t precision);
/* F2C-Filter (BoundaryConditions.txt:69; selected at Lattice.cpp:5179-5211):
 * NOFI hands each coarse node its coincident fine node's non-equilibrium, LAG
 * the mean non-equilibrium of that fine node and its 18 D3Q19 neighbours
 * (Lagrava et al. 2012), TOU an anisotropic filter (Touil et al. 2014): each
 * population's non-equilibrium averaged along its own lattice direction with
 * weights 1/4, 1/2, 1/4 (the rest population unfiltered), cleared of the mass
 * and momentum it carries so that the coarse node keeps the fine node's
 * density and velocity. */
enum { ACP_F2C_NOFI = 0, ACP_F2C_LAG = 1, ACP_F2C_TOU = 2 };
int acp_refined_set_filter(acp_refined* r, int32_t filter);
/* Scaling-Style (BoundaryConditions.txt:89; selected at Lattice.cpp:4634-4704
 * between interpolTimeScalingCoarseToFine and its _MRT_GSbasis twin,
 * GridCoupling.h:30-36) of the standard coupling: the non-equilibrium part
 * handed between the levels is scaled by one factor, a = tau_f / (2 tau_c)
 * coarse to fine and 1/a fine to coarse (ACP_SCALING_SRT, SRTSTYLE, the
 * default), or moment by moment in the Gram-Schmidt basis of the MRT-GS
 * operator, each moment by its own relaxation-time ratio (ACP_SCALING_MRT,
 * MRTSTYLE): the moments that operator relaxes with omega by a, those it
 * relaxes at rate 1 on both levels by 1/2 (2 fine to coarse), the conserved
 * ones by 0.  Applies to the coarse-to-fine rescaling and to every
 * fine-to-coarse filter; the direct coupling has no moment-space form and
 * ignores it.  MRTSTYLE needs the coupling arithmetic in float64. */
enum { ACP_SCALING_SRT = 0, ACP_SCALING_MRT = 1 };
int acp_refined_set_scaling(acp_refined* r, int32_t style);
/* Coupling-Approach (BoundaryConditions.txt:52-53, 71-83): the standard
 * coupling (STD, the default; every link replaced: Coarse-to-Fine fAll,
 * Fine-to-Coarse cAll, Hanging-Node-Reconstruction hAll) or the direct
 * coupling of Astoul et al. 2021 (DC1 / DC2; the reference orchestrates it at
 * Lattice.cpp:13599-15050 with interpolTimeDirectCoupling and directCoupling,
 * GridCoupling.h:28-31, whose arithmetic it does not contain).  Under DC the
 * coarse interface columns x0 (and x1) and the fine face columns coincide and
 * exchange only links: the fine node takes the links arriving from outside
 * the fine region (ACP_LINKS_MISSING, Mf) or those and the ones in the face
 * plane (ACP_LINKS_MISSING_COMMON, MfP) from the coarse node, time-
 * interpolated at t + 1/2; at t + 1 the coarse node takes the links arriving
 * from inside (Mc) or those and the in-plane ones (McP) from the fine node;
 * not MfP with McP.  Non-equilibrium parts are rescaled by a = tau_f /
 * (2 tau_c) about the pair's equilibrium, whose density and velocity are
 * those of the coarse node completed with the rescaled fine-side links.
 * Hanging fine face nodes average their coincident neighbours' links of
 * `hanging` (hMf ACP_LINKS_MISSING, hMfP ACP_LINKS_MISSING_COMMON, hAll
 * ACP_LINKS_ALL).  The standard coupling takes link sets too
 * (BoundaryConditions.txt:73-86): the fine face nodes at coarse positions
 * take from the coarse node the links of `c2f` (Mf, MfP or fAll) and keep what
 * the fine step pushed on the others, the hanging face nodes average those of
 * `hanging`, and with Fine-to-Coarse Mc (ACP_LINKS_MISSING) the coarse
 * grid-coupling columns x0 + 1 and x1 - 1 take from the fine level only their
 * links arriving from inside the fine region (cAll: every link; McP belongs
 * to the direct coupling).  The other restricted coarse nodes take every
 * link.  Every operator and tier; HRR hands its strain rates as
 * the reference's DirectCoupling function computes them (Node.cpp:1892-2111):
 * the coarse interface column and its fine partners collide with the coarse
 * stencil that reads the missing inner neighbour through the fine level, the
 * fine face nodes at t + 1/2 with the coarse-spaced stencil over fine nodes
 * and the time-averaged coarse neighbour outside. */
enum { ACP_COUPLING_STD = 0, ACP_COUPLING_DC = 1 };
enum { ACP_LINKS_MISSING = 0, ACP_LINKS_MISSING_COMMON = 1, ACP_LINKS_ALL = 2 };
int acp_refined_set_coupling(acp_refined* r, int32_t scheme, int32_t c2f, int32_t f2c, int32_t hanging);
/* The fine-to-coarse restriction alone, on the levels' current states (the
 * last part of acp_refined_step; for tests and for a caller that edits a
 * level's state). */
int acp_refined_restrict(acp_refined* r);
/* Placement and graphs, as acp_case_footprint / acp_case_create_in /
 * acp_case_enable_graphs for a single case (no counterpart in the reference):
 * the bytes a refined case needs with `nprobe` probes of `probe_capacity`
 * coarse steps on its coarse level; creation inside a caller's 256-byte
 * aligned device region of at least that size, on the caller's stream (fine
 * level first, coarse level after it); and captured graphs of `chunk` coarse
 * steps (even; 0 turns them off). */
int acp_refined_footprint(const acp_case_desc* coarse, int32_t x0, int32_t x1, int32_t nprobe, int32_t probe_capacity,
                          int64_t* bytes);
int acp_refined_create_in(const acp_case_desc* coarse, int32_t x0, int32_t x1, void* stream, void* region,
                          int64_t region_bytes, acp_refined** out);
int acp_refined_enable_graphs(acp_refined* r, int32_t chunk);

/* ------------------------------------------------------------------ packed stepping
 * No counterpart in the reference, which runs one case per process
 * (main.cpp:37-271).  A pack advances several resident one-level cases of one
 * storage tier and equilibrium order with shared kernel launches: per step
 * the HRR pre-pass, one bulk collide-and-push launch per operator present,
 * one launch per LODI phase, one probe launch and one bookkeeping launch,
 * each a persistent grid walking a member table held in device memory.
 * Members join (acp_pack_add) and leave (acp_pack_remove) between steps; each
 * takes exactly the steps it was given, and its populations, probe series and
 * counters are bit-identical to stepping it alone with acp_case_step.
 * Eligible cases: LODI on every face (ACP_BC_LODI), no cubic-Mach correction,
 * not a level of a two-level case, at most 32 members.  While a case is in a
 * pack, its own stepping, probe set-up and read-back calls return
 * ACP_E_STATE. */
typedef struct acp_pack acp_pack;
/* precision ACP_PREC_*, order 2..4; stream NULL: the pack creates its own */
int acp_pack_create(int32_t precision, int32_t order, void* stream, acp_pack** out);
/* Waits for the pack's work, removes every member (as acp_pack_remove) and
 * frees the pack. */
void acp_pack_destroy(acp_pack* p);
/* Join with `steps` steps to take (initialised case; waits for its stream). */
int acp_pack_add(acp_pack* p, acp_case* c, int32_t steps);
/* Launch n packed steps on the pack's stream; a member takes a step only while
 * it has steps left.  Chunks of acp_pack_enable_graphs steps replay a
 * captured graph. */
int acp_pack_step(acp_pack* p, int32_t n);
int acp_pack_enable_graphs(acp_pack* p, int32_t chunk);
/* Steps of a member not yet launched. */
int acp_pack_steps_left(const acp_pack* p, const acp_case* c, int32_t* left);
/* Leave the pack (waits for the launches that stepped it): the case's arrays,
 * step and probe counters are its own again. */
int acp_pack_remove(acp_pack* p, acp_case* c);
int acp_pack_members(const acp_pack* p, int32_t* members);

/* ------------------------------------------------------------------ serving
 * No counterpart in the reference, which runs one case per process
 * (main.cpp:37-271).  A server keeps many independent cases resident on one
 * GPU and steps them concurrently:
 *   - requests are admitted in arrival order once they have arrived, one of
 *     `max_active` stream slots is free and their footprint
 *     (acp_case_footprint) fits a free range of one device arena of
 *     `hbm_budget_bytes` allocated when the server starts;
 *   - each resident case runs on its own stream in chunks of `graph_chunk`
 *     steps replayed from captured HIP graphs (or `launch_steps` plain steps
 *     when graph_chunk is 0), with at most `chunks_ahead` chunks queued;
 *   - with `pack_max_nodes` set, eligible one-level cases of at most that
 *     many nodes are stepped together instead (acp_pack_*): one pack per tier
 *     and order, each on its own stream, chunks of `graph_chunk` steps.  Small
 *     cases are bound by launch latency, which packing removes; large ones
 *     run faster on their own streams, where their kernels overlap;
 *   - a finished case posts an acp_result and gives back its range and slot.
 * Results are those of the same case run alone: cases share no state. */
enum { ACP_INIT_REFERENCE = 0, ACP_INIT_PULSE = 1, ACP_INIT_RANDOM = 2 };

typedef struct acp_server_desc {
    int64_t hbm_budget_bytes;   /* device arena for resident cases */
    int32_t max_active;         /* resident cases (one stream each) */
    int32_t graph_chunk;        /* steps per captured graph replay (even), 0: plain launches */
    int32_t launch_steps;       /* steps per acp_case_step call when graph_chunk is 0 */
    int32_t chunks_ahead;       /* chunks queued per case before the driver waits on it */
    int32_t admit_threads;      /* host threads creating and initialising admitted cases (0: 4) */
    int32_t pack_max_nodes;     /* > 0: eligible one-level cases (acp_pack_add) of at most this many nodes
                                   share packed step launches, one pack per tier and order; 0: every case
                                   on its own stream */
} acp_server_desc;

typedef struct acp_request {
    acp_case_desc desc;
    int32_t steps;
    int32_t init_kind;          /* ACP_INIT_*: acp_case_init_reference / _pulse / _random */
    double amp, x0, y0, halfwidth;      /* pulse parameters (acp_case_init_pulse) */
    uint64_t seed;
    double rho_amp, u_amp;              /* random field (acp_case_init_random) */
    int32_t probe[3];           /* probe node (i, j, k); probe[0] < 0: none */
    double* probe_out;          /* host buffer for `steps` densities, or NULL; must stay valid */
    double arrival;             /* seconds on the server clock; not admitted earlier */
    uint64_t tag;               /* caller's identifier, returned in the result */
    int32_t refine[2];          /* fine region [x0, x1] of a two-level case (acp_refined_*; periodic or
                                   LODI_X coarse level, pulse initial condition); refine[1] <= refine[0]:
                                   one level */
    int32_t f2c_filter;         /* ACP_F2C_* of a two-level case */
    int32_t coupling;           /* ACP_COUPLING_* of a two-level case; the direct coupling exchanges the
                                   missing links (Mf, Mc, hMf) */
} acp_request;

typedef struct acp_result {
    uint64_t tag;
    int32_t status;             /* ACP_OK or the error of the failing call */
    int32_t steps;              /* steps launched */
    int64_t device_bytes;       /* footprint the case held */
    double t_arrival, t_submit, t_admit, t_start, t_done;  /* server clock, seconds */
    double mass;                /* acp_case_mass of the final state (coarse level of a two-level case) */
    uint64_t checksum;          /* acp_case_checksum of the final state (two levels: coarse + fine, mod 2^64) */
    char error[128];
} acp_result;

typedef struct acp_server_stats {
    int64_t submitted, completed, queued;
    int32_t active, peak_active;
    int64_t bytes_in_use, peak_bytes, budget_bytes;
    double clock;
} acp_server_stats;

typedef struct acp_server acp_server;

/* Start a server on the current device (allocates the arena and streams and
 * starts its admission and driver threads). */
int acp_server_create(const acp_server_desc* desc, acp_server** out);
/* Queue a request (copied); thread-safe.  Fails at once when the case is
 * invalid or larger than the whole budget. */
int acp_server_submit(acp_server* s, const acp_request* r);
/* Move up to `max` finished results to `out`, waiting up to `wait_ms` for one. */
int acp_server_poll(acp_server* s, acp_result* out, int32_t max, int32_t wait_ms, int32_t* n);
int acp_server_get_stats(acp_server* s, acp_server_stats* out);
/* Seconds since acp_server_create (the clock of arrival and result times). */
double acp_server_clock(acp_server* s);
/* Finish every submitted request, then stop the threads and free the arena. */
void acp_server_destroy(acp_server* s);

/* Measurement helper, no counterpart in the reference: the device-copy
 * ceiling for a population layout.  Copies `planes` (1 or 19) float64 planes
 * of `plane_elems` elements each, one element of every plane per thread (the
 * access pattern of the bulk kernel without the neighbour shift); with 19
 * planes and `block_nodes` > 0 (a power of two) the planes are laid out in
 * blocks of that many nodes as the engine lays out its float64 and float32
 * populations, with 0 as whole planes.  Like the bulk kernels, the copy
 * carries the non-temporal hint when its two buffers exceed the 256 MB
 * last-level cache.  `reps` copies on `stream` after one untimed copy; writes
 * the mean milliseconds per copy.  Allocates and frees its own two buffers. */
int acp_copy_probe(int32_t planes, int64_t plane_elems, int64_t block_nodes, int32_t reps, void* stream,
                   double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
