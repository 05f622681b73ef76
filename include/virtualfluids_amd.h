/*
 * virtualfluids_amd — C ABI of the MI355X (gfx950) LBM time step.
 *
 * One GPU owns the whole domain. Every entry point launches asynchronously on the
 * given HIP stream (pass the hipStream_t as `stream`; NULL is the null stream),
 * never allocates device memory per call, and returns VF_OK or an error code;
 * vf_get_last_error() then describes the failure. All pointers inside the
 * structs are device pointers owned by the caller, unless marked "host".
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the VirtualFluids source tree).
 */
#ifndef VIRTUALFLUIDS_AMD_H
#define VIRTUALFLUIDS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VF_ABI_VERSION 6

enum vf_status {
    VF_OK = 0,
    VF_ERR_INVALID_ARGUMENT = 1, /* null pointer, size out of range, unknown enum value */
    VF_ERR_HIP = 2,              /* the HIP runtime reported an error on launch */
    VF_ERR_UNSUPPORTED = 3       /* valid request this build does not implement */
};

/* Storage type of the population buffer. Arithmetic is always FP32. */
enum vf_storage {
    VF_STORAGE_FP32 = 0,
    VF_STORAGE_FP16 = 1, /* IEEE binary16 shifted populations f_i - w_i */
    VF_STORAGE_BF16 = 2  /* bfloat16 shifted populations f_i - w_i */
};

/* CollisionTemplate (src/gpu/core/Calculation/Calculation.h:56-69) */
enum vf_collision_template {
    VF_COLLISION_DEFAULT = 0,
    VF_COLLISION_WRITE_MACRO_VARS = 1,
    VF_COLLISION_APPLY_BODY_FORCE = 2,
    VF_COLLISION_ALL_FEATURES = 3,
    VF_COLLISION_SUBDOMAIN_BORDER = 4
};

/* Node types (src/gpu/core/Calculation/Calculation.h:36-45) */
#define VF_GEO_PM_0 5u
#define VF_GEO_PM_1 6u
#define VF_GEO_PM_2 7u
#define VF_GEO_SOLID 15u
#define VF_GEO_VOID 16u
#define VF_GEO_FLUID 19u

/*
 * Kernel choice of the dense-box collision (vf_k17_run_box, vf_k17_run_box_into) and its adjoint
 * (vf_k17_adjoint_box). No counterpart in the reference. All fields zero (the default of a
 * zero-initialised vf_lbm_parameter) selects automatically, by row length, storage type and
 * workingSetBytes; any other value pins one form, for benchmarks and tests. A pinned form that
 * does not fit the box (a row width that does not divide nx, or rows longer than one workgroup)
 * falls back to the one-node-per-thread kernel. The results agree to FP32 round-off.
 */
enum vf_box_kernel {
    VF_BOX_KERNEL_AUTO = 0,
    VF_BOX_KERNEL_ROW = 1,  /* one workgroup per x-row (streaming form) */
    VF_BOX_KERNEL_NODE = 2  /* one node per thread (cache-resident form) */
};
enum vf_row_addressing {
    VF_ROW_ADDRESSING_AUTO = 0,   /* buffer descriptors where 14 direction slots fit in 4 GiB */
    VF_ROW_ADDRESSING_BUFFER = 1, /* buffer descriptors (only where they fit; else 64-bit) */
    VF_ROW_ADDRESSING_GLOBAL = 2  /* 64-bit global addresses */
};
enum vf_adjoint_kernel {
    VF_ADJOINT_KERNEL_AUTO = 0,
    VF_ADJOINT_KERNEL_ROW_WHOLE = 1, /* row kernel over whole rows of any length (up to 512) */
    VF_ADJOINT_KERNEL_NODE = 2,      /* one node per thread */
    VF_ADJOINT_KERNEL_ROW = 3        /* row kernel over the whole waves of every row, and a launch of
                                        the one-node-per-thread kernel on the columns beyond where
                                        the last wave would carry fewer than 32 lanes */
};
typedef struct vf_kernel_choice {
    int32_t boxKernel;          /* enum vf_box_kernel */
    int32_t rowWidth;           /* 0: automatic; 1, 2 or 4 nodes per lane (2 and 4 for FP16/BF16) */
    int32_t rowWorkgroupsPerCU; /* 0: automatic; -1: no cap; k > 0: at most k row workgroups per CU */
    int32_t rowAddressing;      /* enum vf_row_addressing (collision and adjoint row kernels) */
    int32_t rowXcdRanges;       /* 0: automatic; 1: contiguous row ranges per XCD; 2: launch order */
    int32_t adjointKernel;      /* enum vf_adjoint_kernel */
} vf_kernel_choice;

/*
 * One grid level: the fields of LBMSimulationParameter the time step reads
 * (src/gpu/core/Parameter/Parameter.h:67-256).
 *
 * Populations: ONE in-place buffer of 27*numberOfNodes values, structure of arrays
 * f[d*numberOfNodes + k], d in the D3Q27 order of src/lbm/constants/D3Q27.h:46-72
 * (src/gpu/core/Cuda/CudaMemoryManager.cpp:169). Esoteric-twist parity
 * isEvenTimestep selects the slot table (src/gpu/core/Utilities/KernelUtilities.h:66-125).
 *
 * Addressing: neighborX/Y/Z hold the +x/+y/+z neighbour of every node
 * (Parameter.h:84-86). When nx*ny*nz == numberOfNodes the node index is
 * k = x + nx*(y + ny*z) of a periodic box and vf_k17_run_box() recomputes the
 * neighbours instead of loading them; every other entry point except the indexed
 * collision (vf_k17_run, vf_k17_run_on_indices) and its adjoint (vf_k17_adjoint) also takes
 * NULL neighbour arrays for such a box and then recomputes them too.
 */
typedef struct vf_lbm_parameter {
    void* distributions;            /* 27*numberOfNodes values of type `storage` */
    int32_t storage;                /* enum vf_storage */
    int32_t isEvenTimestep;         /* Parameter.h:71 */
    uint64_t numberOfNodes;         /* includes any dummy node 0 of a sparse grid */
    const uint32_t* neighborX;      /* +x neighbour, numberOfNodes entries */
    const uint32_t* neighborY;      /* +y neighbour */
    const uint32_t* neighborZ;      /* +z neighbour */
    const uint32_t* typeOfGridNode; /* Parameter.h:79, VF_GEO_* */
    int32_t nx, ny, nz;             /* dense periodic box extents (0 when sparse) */
    const uint8_t* fluidMask;       /* box path only: collide where != 0 (NULL: all nodes) */
    float omega;                    /* 1/(3 nu + 1/2) on this level (Parameter.cpp:400-418) */
    float quadricLimiters[3];       /* host values P, M, D (Parameter.cpp:69, default 0.01) */
    float forces[3];                /* host values, global body force (Parameter::getForcesDev) */
    float forceFactor;              /* 2^-level (src/gpu/core/Utilities/KernelUtilities.h:56-64) */
    float* bodyForceX;              /* per-node force, ApplyBodyForce templates (reset to 0 after use) */
    float* bodyForceY;
    float* bodyForceZ;
    float* rho;                     /* drho = rho - 1 output / init input */
    float* velocityX;
    float* velocityY;
    float* velocityZ;
    float* pressure;
    const uint32_t* fluidNodeIndices; /* taggedFluidNodeIndices[Default] (Parameter.h:245-256) */
    uint32_t numberOfFluidNodes;
    /* Brinkman porosity per node (NULL: none; no counterpart in the reference): a drag
     * acceleration -porosity * u applied through the collision's half-step velocity shift,
     * so the collision velocity becomes u (1 - porosity/2). Honoured by every collision and
     * adjoint entry point. */
    const float* porosity;
    /* Bytes of lattice state that the caller's time step cycles through between two collisions
     * of this lattice, e.g. every level of a refinement hierarchy (0: this lattice's own
     * 27 * numberOfNodes values; no counterpart in the reference). Dense-box collisions pick
     * their kernel from it: up to the 256 MB infinity cache of an MI355X the work is
     * cache-resident and latency-bound and takes the one-node-per-thread kernel, beyond it the
     * streaming row kernel. */
    uint64_t workingSetBytes;
    /* Kernel choice of the dense-box collision and its adjoint (zero: automatic; ABI 4). */
    vf_kernel_choice kernelChoice;
} vf_lbm_parameter;

/*
 * QforBoundaryConditions / QforDirectionalBoundaryCondition
 * (src/gpu/core/Calculation/Calculation.h:105-127).
 * q27 is SoA: q27[d*numberOfBCnodes + i]; a link is active when 0 <= q <= 1.
 */
typedef struct vf_q_boundary_conditions {
    const int32_t* k;         /* BC node indices */
    const int32_t* kN;        /* pressure BC: neighbour node whose state is extrapolated */
    const float* q27;         /* 27*numberOfBCnodes sub-grid distances */
    uint32_t numberOfBCnodes;
    const float* Vx;          /* velocity BCs: wall velocity per BC node */
    const float* Vy;
    const float* Vz;
    const float* RhoBC;       /* pressure BC: prescribed drho per BC node */
    int32_t direction;        /* pressure BC: D3Q27 index of the outflow side (dP00, dM00, ...) */
    /* Optional (NULL: unused; no counterpart in the reference; ABI 5): 27*numberOfBCnodes floats,
     * postCollision[d*numberOfBCnodes + i]. The post-collision BC kernels (vf_NoSlip*, vf_Velocity*)
     * store there the 27 post-collision values each BC node's links were computed from; their
     * adjoints then read the linearisation point from it instead of recomputing each BC node's
     * collision. */
    float* postCollision;
    /* Layout hint for the adjoint kernels (no counterpart in the reference; ABI 6): nonzero when
     * consecutive entries of k are, for the most part, neighbours in x (walls normal to y or z).
     * The adjoints then give each BC node one thread, so that the loads of a direction are
     * contiguous; zero gives each node a lane per direction, which keeps the scattered cells of a
     * face normal to x in flight together. Either value gives the same adjoint up to the order of
     * the per-node sums. */
    int32_t nodesAlongRows;
} vf_q_boundary_conditions;

/* ICells / ICellNeigh (src/gpu/core/Calculation/Calculation.h:72-89) */
typedef struct vf_interpolation_cells {
    const uint32_t* fineCellIndices;
    const uint32_t* coarseCellIndices;
    uint32_t numberOfCells;
} vf_interpolation_cells;

typedef struct vf_interpolation_cell_neighbor {
    const float* x; /* per-cell offset of the interpolation point (refinement towards walls) */
    const float* y;
    const float* z;
} vf_interpolation_cell_neighbor;

/* ---------------------------------------------------------------------------------
 * Library
 * --------------------------------------------------------------------------------- */
int vf_abi_version(void);
const char* vf_get_last_error(void);

/* Device-to-device copy of `bytes` (a multiple of 16, both pointers 16-byte aligned) with
 * aligned non-temporal 16-byte accesses. No counterpart in the reference: the copy kernel whose
 * bandwidth the collision kernels are measured against. */
int vf_stream_copy(const void* src, void* dst, uint64_t bytes, void* stream);

/* In-place read-modify-write of 27 streams of streamBytes each (stream d at buf + d*streamBytes;
 * streamBytes a multiple of 1024 and below 2^34): every element is read and written into the
 * opposite stream at the same offset (d <-> inv(d) of the D3Q27 order, stream 0 in place), the
 * memory pattern of the in-place collision step without its arithmetic (27 reads and 27 writes
 * per node), in workgroups of 256 elements per stream; workgroupsPerCU (0: no cap, else 1-32)
 * caps the resident workgroups per CU as the FP32 row kernel's residency cap does. No
 * counterpart in the reference: the yardstick the collision kernels' bandwidth is measured
 * against. */
int vf_stream_inplace27(void* buf, uint64_t streamBytes, int32_t workgroupsPerCU, void* stream);

/* ---------------------------------------------------------------------------------
 * K17 collide + eso-twist stream
 * --------------------------------------------------------------------------------- */

/* K17CompressibleNavierStokes<None>::run() (src/gpu/core/Kernel/Compressible/NavierStokes/K17/K17CompressibleNavierStokes.cu:83-101):
 * every node of fluidNodeIndices, CollisionTemplate::Default. */
int vf_k17_run(const vf_lbm_parameter* para, void* stream);

/* K17CompressibleNavierStokes<None>::runOnIndices(indices, n, template, stream)
 * (K17CompressibleNavierStokes.cu:103-145; kernel body src/gpu/core/Utilities/RunCollision.cuh:74-143). */
int vf_k17_run_on_indices(const vf_lbm_parameter* para, const uint32_t* indices, uint32_t numberOfIndices,
                          int32_t collisionTemplate, void* stream);

/* The same update on a dense periodic box (nx, ny, nz set): neighbours recomputed from
 * the node coordinates, no index traffic. Replaces run()/runOnIndices() for uniform
 * boxes; optional fluidMask skips solid nodes (NoSlipDelayBounceBack). */
int vf_k17_run_box(const vf_lbm_parameter* para, int32_t collisionTemplate, void* stream);

/* The same step out of place: reads para->distributions, writes the stepped populations to
 * `out` (27*numberOfNodes values of the same storage type, a different buffer); masked nodes'
 * cells are copied unchanged, so `out` equals what the in-place step leaves. No counterpart in
 * the reference: it lets the adjoint's replay keep every intermediate state without copies. */
int vf_k17_run_box_into(const vf_lbm_parameter* para, void* out, void* stream);

/* Post-collision bounce-back lists that the dense-box collision applies in the same launch
 * (vf_k17_run_box_bounce_back; no counterpart in the reference). lists[l] are BC lists in the
 * order their separate launches would run (NULL where unused); velocity[l] selects
 * VelocityBounceBack (1: the list needs Vx, Vy, Vz) or NoSlipBounceBack (0). nodeEntry (device,
 * numberOfNodes values) holds -1 or the node's entry; per entry, entryIndex (device, 3 values)
 * holds the node's index in list l or -1, and entryLinks (device, 3 values) the bitmask of the
 * links d with 0 <= q_d <= 1 the node has in list l. Every such link must lead to a node the
 * collision does not update (fluidMask 0, as at box walls): the epilogue then writes exactly what
 * the separate kernels would. */
typedef struct vf_post_bounce_back {
    const vf_q_boundary_conditions* lists[3]; /* host pointers */
    int32_t velocity[3];
    const int32_t* nodeEntry;
    const int32_t* entryIndex;
    const uint32_t* entryLinks;
    uint32_t numberOfEntries;
} vf_post_bounce_back;

/* vf_k17_run_box followed by the bounce-back lists of bb in their order (VelocityBounceBack or
 * NoSlipBounceBack, Velocity.cu:44-63 / NoSlip.cu:44-60), with identical results bit for bit.
 * Where the one-node-per-thread kernel serves the box it writes the bounce-back values in its own
 * epilogue, from the post-collision values it holds, so the BC launches and their reads of the
 * post-collision state disappear; where the row kernel serves it, the BC kernels run after it. */
int vf_k17_run_box_bounce_back(const vf_lbm_parameter* para, int32_t collisionTemplate, const vf_post_bounce_back* bb,
                               void* stream);

/* ---------------------------------------------------------------------------------
 * Boundary conditions: the free launchers BoundaryConditionFactory hands out
 * (src/gpu/core/BoundaryConditions/BoundaryConditionFactory.cpp:49-217).
 * Post-collision kernels read with the current parity and write the flipped one.
 * --------------------------------------------------------------------------------- */

/* NoSlipBounceBack (src/gpu/core/BoundaryConditions/NoSlip/NoSlip.cu:44-60) */
int vf_NoSlipBounceBack(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, void* stream);
/* NoSlipInterpolatedCompressible (NoSlip.cu:81-98) */
int vf_NoSlipInterpolatedCompressible(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, void* stream);
/* VelocityBounceBack (src/gpu/core/BoundaryConditions/Velocity/Velocity.cu:44-63) */
int vf_VelocityBounceBack(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, void* stream);
/* VelocityInterpolatedCompressible (Velocity.cu:87-107) */
int vf_VelocityInterpolatedCompressible(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, void* stream);
/* VelocityWithPressureInterpolatedCompressible (Velocity.cu:109-129) */
int vf_VelocityWithPressureInterpolatedCompressible(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc,
                                                    void* stream);
/* PressureNonEquilibriumCompressible, pre-collision, after the parity swap
 * (src/gpu/core/BoundaryConditions/Pressure/Pressure.cu:63-81) */
int vf_PressureNonEquilibriumCompressible(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc,
                                          void* stream);

/* ---------------------------------------------------------------------------------
 * Adjoint of the collide + stream step (no counterpart in the reference: the reverse mode
 * behind the differentiable time step of the Python layer).
 *
 * para->distributions holds the populations BEFORE the forward step and para->isEvenTimestep
 * the parity that step used (the forward step's linearisation point). adjoint is an FP32
 * buffer of 27*numberOfNodes values in the same slot layout: on entry the adjoint of the
 * populations after the step, on exit (in place) the adjoint of the populations before it.
 * velocityShiftAdjoint (optional, 3*numberOfNodes floats: all x, then all y, then all z): the
 * adjoint of each node's collision velocity shift (forces * 1/2 * forceFactor) is ADDED, so
 * per-node force or porosity gradients accumulate over steps. Nodes that the forward step does
 * not collide (fluidMask == 0, or not listed) are left untouched, as in the forward step.
 * --------------------------------------------------------------------------------- */

/* Reverse of vf_k17_run_box (dense periodic box, optional fluidMask). porosityAdjoint
 * (optional, numberOfNodes floats): with para->porosity set, the adjoint of each node's
 * porosity is ADDED. */
int vf_k17_adjoint_box(const vf_lbm_parameter* para, float* adjoint, float* velocityShiftAdjoint,
                       float* porosityAdjoint, void* stream);

/* Reverse of vf_k17_run (every node of fluidNodeIndices, neighbour arrays). */
int vf_k17_adjoint(const vf_lbm_parameter* para, float* adjoint, float* velocityShiftAdjoint, float* porosityAdjoint,
                   void* stream);

/* Adjoints of the BC kernels, run in reverse order within a step (pressure BC first, then the
 * post-collision BCs in reverse launch order, then the collision adjoint). For the post BCs
 * para->distributions holds the populations BEFORE the step (each BC node's collision is
 * recomputed from them with para's omega, limiters and force), or bc->postCollision holds the
 * post-collision values the forward BC kernel stored (then nothing is recomputed); for the
 * pressure BC para->distributions holds the populations after the step. para->isEvenTimestep is the parity the forward kernel used.
 * The adjoint of every written cell moves to the cells the value was computed from and the
 * written cell's adjoint is cleared. velocityAdjoint (optional, 3*numberOfBCnodes floats: Vx,
 * Vy, Vz blocks) and rhoBCAdjoint (optional, numberOfBCnodes) accumulate parameter gradients.
 * No counterpart in the reference. */
int vf_NoSlipBounceBackAdjoint(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, float* adjoint,
                               float* velocityAdjoint, void* stream);
int vf_NoSlipInterpolatedCompressibleAdjoint(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc,
                                             float* adjoint, float* velocityAdjoint, void* stream);
int vf_VelocityBounceBackAdjoint(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, float* adjoint,
                                 float* velocityAdjoint, void* stream);
int vf_VelocityInterpolatedCompressibleAdjoint(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc,
                                               float* adjoint, float* velocityAdjoint, void* stream);
int vf_VelocityWithPressureInterpolatedCompressibleAdjoint(const vf_lbm_parameter* para,
                                                           const vf_q_boundary_conditions* bc, float* adjoint,
                                                           float* velocityAdjoint, void* stream);
/* neighborsDistinct: nonzero when no two nodes of the list share a kN (as for the inward
 * neighbours of one face); each neighbour's cells are then updated by one thread with plain
 * read-modify-writes, else with atomic adds. */
int vf_PressureNonEquilibriumCompressibleAdjoint(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc,
                                                 float* adjoint, float* rhoBCAdjoint, int32_t neighborsDistinct,
                                                 void* stream);

/* ---------------------------------------------------------------------------------
 * Grid refinement: the launchers GridScalingFactory hands out for
 * GridScaling::ScaleCompressible (src/gpu/core/GridScaling/GridScalingFactory.cpp:43-69),
 * whose signature is gridScaling(LBMSimulationParameter* coarse, LBMSimulationParameter* fine,
 * ICells*, ICellNeigh&, stream) (src/gpu/core/GridScaling/GridScalingFactory.h:46). Coarse and fine levels must share the
 * storage type; the coarse level's isEvenTimestep selects the coarse parity, the fine level is
 * read/written with the even parity (it has just completed two steps).
 * offsets may be NULL (all cells centred) or hold per-cell x/y/z offsets.
 * --------------------------------------------------------------------------------- */

/* scaleCoarseToFineCompressible<false> (src/gpu/core/GridScaling/Scaling.cu:105-138,
 * kernel src/gpu/core/GridScaling/scaleCoarseToFineCompressible.cu:305-528): for every cell
 * i, the 8 coarse nodes of the cell at coarseCellIndices[i] give the interpolation
 * polynomials; the 8 fine nodes of the cell at fineCellIndices[i] are overwritten. */
int vf_ScaleCoarseToFineCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                     const vf_interpolation_cells* coarseToFine,
                                     const vf_interpolation_cell_neighbor* offsets, void* stream);

/* scaleFineToCoarseCompressible<false> (Scaling.cu:155-188, kernel
 * src/gpu/core/GridScaling/scaleFineToCoarseCompressible.cu:76-129): the 8 fine nodes of the
 * cell at fineCellIndices[i] give the polynomials; coarse node coarseCellIndices[i] is
 * overwritten. */
int vf_ScaleFineToCoarseCompressible(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                     const vf_interpolation_cells* fineToCoarse,
                                     const vf_interpolation_cell_neighbor* offsets, void* stream);

/* ---------------------------------------------------------------------------------
 * Initialisation, output, conservation
 * --------------------------------------------------------------------------------- */

/* InitNavierStokesCompressible::init (src/gpu/core/PreProcessor/PreProcessorStrategy/InitNavierStokesCompressible/InitNavierStokesCompressible.cu:45-87):
 * shifted equilibrium from para->rho (drho) and para->velocity*, every node that is not
 * VF_GEO_SOLID/VF_GEO_VOID (all nodes when typeOfGridNode is NULL). */
int vf_InitNavierStokesCompressible(const vf_lbm_parameter* para, void* stream);

/* InitNavierStokesCompressible::init with para->getUseInitNeq() (InitNavierStokesCompressible.cu:67-84,
 * kernel InitNavierStokesCompressibleNonEquilibrium_Device, InitNavierStokesCompressible_Device.cu:216-501):
 * the equilibrium plus (1 + drho) times the first-order non-equilibrium part from central
 * differences of velocityX/Y/Z, relaxation rate para->omega. Solid and void nodes get 96 in their
 * rest slot (the reference's marker). neighborInverse: the (-1,-1,-1) neighbour of every node
 * (Parameter.h neighborInverse); may be NULL for a dense periodic box without neighbour arrays. */
int vf_InitNavierStokesCompressibleNonEquilibrium(const vf_lbm_parameter* para, const uint32_t* neighborInverse,
                                                  void* stream);

/* calculateMacroscopicQuantitiesCompressible (src/gpu/core/PostProcessor/MacroscopicQuantities.cu:194-237, 953):
 * fills para->rho, velocity*, pressure from the pre-collision state. */
int vf_CalculateMacroscopicQuantitiesCompressible(const vf_lbm_parameter* para, void* stream);

/* FP64 conservation sums over the pre-collision state of every node with
 * typeOfGridNode == VF_GEO_FLUID (all nodes when NULL):
 * out4[0] = sum drho, out4[1..3] = sum of momentum j. out4 is a device pointer to 4 doubles,
 * overwritten. No counterpart in the reference (new: mass/momentum checks). */
int vf_ConservationSums(const vf_lbm_parameter* para, double* out4, void* stream);

/* One FP64 pass over the same nodes: out5[0..3] as vf_ConservationSums, out5[4] = sum of the
 * kinetic energy 1/2 (1+drho)|u|^2 = |j|^2 / (2 (1+drho)) (out5 device pointer, overwritten).
 * The energy term is kineticEnergyKernel's (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:99-153)
 * over every fluid node; vf_KineticEnergy applies the reference's node selection. */
int vf_MomentSums(const vf_lbm_parameter* para, double* out5, void* stream);

/* Adjoint of vf_ScaleCoarseToFineCompressible (no counterpart in the reference: the reverse mode
 * of the refinement interface for gradients through two levels). coarseAdjoint / fineAdjoint:
 * FP32 adjoint buffers with the population layout of each level. The adjoint of the fine cells
 * the forward call wrote is read and cleared; its transpose through the interpolation is added
 * to the coarse source nodes' cells (coarse parity). Linearisation point: the coarse state the
 * forward call read (same levels, parities and cells). */
int vf_ScaleCoarseToFineCompressibleAdjoint(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                            const vf_interpolation_cells* coarseToFine,
                                            const vf_interpolation_cell_neighbor* offsets, float* coarseAdjoint,
                                            float* fineAdjoint, void* stream);

/* The distinct source nodes of a set of interface cells: node[u] for u < numberOfNodes, and
 * first[u] .. first[u + 1] (numberOfNodes + 1 entries) indexing lanes, the entries 8 i + q of the
 * cells i whose corner q is node u (every 8 i + q once; for coarse-to-fine cells the corners of
 * the coarse cell at coarseCellIndices[i], bit 0 of q the +x, bit 1 the +y, bit 2 the +z step). */
typedef struct vf_interpolation_sources {
    const uint32_t* node;
    const uint32_t* first;
    const uint32_t* lanes;
    uint32_t numberOfNodes;
} vf_interpolation_sources;

/* vf_ScaleCoarseToFineCompressibleAdjoint with the transposed interpolation gathered per source
 * node (no counterpart in the reference): the cell kernel stores each corner's moment adjoint in
 * `workspace` (9 x 8 x numberOfCells floats, overwritten), then one thread per source node sums
 * its corners' (in the order of `lanes`) and adds the result to the node's cells, instead of
 * an atomic add per cell corner. The same results up to summation order, deterministic from
 * call to call. */
int vf_ScaleCoarseToFineCompressibleAdjointGather(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                  const vf_interpolation_cells* coarseToFine,
                                                  const vf_interpolation_cell_neighbor* offsets,
                                                  const vf_interpolation_sources* sources, float* workspace,
                                                  float* coarseAdjoint, float* fineAdjoint, void* stream);

/* vf_ScaleCoarseToFineCompressibleAdjointGather that allocates nothing and takes no workspace. No
 * counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Its workspace is a device
 * array inside the library, 18 MiB on every GPU it runs on, held until the process ends, which
 * limits a call to 65536 cells (larger calls return VF_ERR_INVALID_ARGUMENT). Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams, other host threads)
 * race, and must be ordered by the caller. Results are bit-identical to the workspace form. */
int vf_ScaleCoarseToFineCompressibleAdjointGatherNoAlloc(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                                         const vf_interpolation_cells* coarseToFine,
                                                         const vf_interpolation_cell_neighbor* offsets,
                                                         const vf_interpolation_sources* sources, float* coarseAdjoint,
                                                         float* fineAdjoint, void* stream);

/* Adjoint of vf_ScaleFineToCoarseCompressible: the adjoint of the coarse cells it wrote is read
 * and cleared, its transpose added to the fine source nodes' cells (fine parity even). */
int vf_ScaleFineToCoarseCompressibleAdjoint(const vf_lbm_parameter* coarse, const vf_lbm_parameter* fine,
                                            const vf_interpolation_cells* fineToCoarse,
                                            const vf_interpolation_cell_neighbor* offsets, float* coarseAdjoint,
                                            float* fineAdjoint, void* stream);

/* DragLiftPostD27 (src/gpu/core/PostProcessor/DragLift.cu:558-566, kernel DragLift.cu:50-302):
 * per BC node i, the post-collision half of the momentum exchange over the links e with
 * q27[e*n + i] in [0, 1]: drag{X,Y,Z}[i] = -sum_e c_e f_e^post (doubles, device arrays of
 * numberOfBCnodes). Call after the collision and before the BC kernels, with the collision's
 * isEvenTimestep (BoundaryConditionKernelManager.cpp:257-278). */
int vf_DragLiftPostD27(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, double* dragX,
                       double* dragY, double* dragZ, void* stream);

/* DragLiftPreD27 (DragLift.cu:569-577, kernel DragLift.cu:304-556): the pre-collision half after
 * the BC kernels and the parity swap, drag{X,Y,Z}[i] = -sum_e c_inv(e) f_inv(e)^pre
 * (BoundaryConditionKernelManager.cpp:149-172). calcDragLift (DragLift.cpp:67-72) sums
 * pre - post over the nodes. */
int vf_DragLiftPreD27(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, double* dragX,
                      double* dragY, double* dragZ, void* stream);

/* The node sum of calcDragLift (DragLift.cpp:67-72), force3[0..2] = sum_i (pre_i - post_i), in
 * one call on the state after the step (after the swap; the post-collision populations that
 * reached the wall are still in the solid nodes' cells then). FP64, one workgroup, fixed
 * summation order (identical results on every call). force3: device pointer to 3 doubles,
 * overwritten. */
int vf_DragLiftSum(const vf_lbm_parameter* para, const vf_q_boundary_conditions* bc, double* force3, void* stream);

/* KineticEnergyAnalyzer::run (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:66-153):
 * out2[0] = sum of 1/2 |u|^2 (1 + drho) over the nodes of type VF_GEO_FLUID whose six face
 * neighbours and (-1,-1,-1) neighbour are fluid, out2[1] = their number (the reference's mean
 * is out2[0] / out2[1]).
 * Reads para->velocityX/Y/Z and para->rho (fill them with
 * vf_CalculateMacroscopicQuantitiesCompressible first) and typeOfGridNode; neighborInverse
 * (the (-1,-1,-1) neighbour) is needed with neighbour arrays. FP64 sums; out2 device pointer,
 * overwritten. */
int vf_KineticEnergy(const vf_lbm_parameter* para, const uint32_t* neighborInverse, double* out2, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VIRTUALFLUIDS_AMD_H */
