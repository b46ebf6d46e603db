/*
 * alien_hip.h — C ABI of the MI355X physics timestep engine.
 *
 * This is the device seam of the engine: it replaces the CUDA seam
 * `_SimulationCudaFacade` of ALIEN (source/EngineGpuKernels/SimulationCudaFacade.cuh:27-139)
 * for the physics path only.  Plain C types, plain pointers and sizes; nothing throws across it.
 * Every function returns 0 on success and a negative ALIEN_E_* code on failure; the message of the
 * last failure on the calling thread is available from alien_last_error().
 *
 * Wire format: the transfer objects mirror ALIEN's POD DataTO (source/EngineGpuKernels/TOs.cuh),
 * restricted to the fields the physics path reads or carries through unchanged.
 */
#ifndef ALIEN_HIP_H
#define ALIEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALIEN_ABI_VERSION 2

/* ---- error codes ---------------------------------------------------------------------------- */
#define ALIEN_OK 0
#define ALIEN_E_ARG (-1)       /* invalid argument / unsupported parameter value                 */
#define ALIEN_E_OOM (-2)       /* device allocation failed (CudaMemoryAllocationException analog) */
#define ALIEN_E_CAPACITY (-3)  /* a caller buffer or an internal bounded list is too small        */
#define ALIEN_E_HIP (-4)       /* HIP runtime error (CHECK_FOR_CUDA_ERROR analog)                 */
#define ALIEN_E_STATE (-5)     /* object in invalid state (e.g. a failed step poisoned the world) */
#define ALIEN_E_BUDGET (-6)    /* serving: job does not fit the HBM budget                        */
#define ALIEN_E_NOTFOUND (-7)  /* serving: unknown job id                                         */

/* ---- constants (source/EngineInterface/EngineConstants.h) ---------------------------------- */
#define ALIEN_MAX_CELL_BONDS 6
#define ALIEN_MAX_COLORS 7

#define ALIEN_MOTION_FLUID 0     /* MotionType_Fluid, Motion.h:41     */
#define ALIEN_MOTION_COLLISION 1 /* MotionType_Collision, Motion.h:42 */

/* ---- transfer objects ----------------------------------------------------------------------- */

/* ConnectionTO (TOs.cuh:30-35): bond to another cell of the same transfer array. */
typedef struct alien_connection_to {
    int32_t cellIndex;       /* index into the cell array of the same call */
    float distance;          /* rest length                                  */
    float angleFromPrevious; /* degrees, the bond angles of a cell sum to 360 */
} alien_connection_to;

/* CellTO (TOs.cuh:156-192), physics subset plus pass-through identity fields. */
typedef struct alien_cell_to {
    uint64_t id;
    alien_connection_to connections[ALIEN_MAX_CELL_BONDS];
    float pos[2];
    float vel[2];
    float energy;
    float stiffness;
    uint8_t color;
    uint8_t maxConnections;
    uint8_t numConnections;
    uint8_t barrier;
    uint32_t age;
    uint8_t livingState; /* LivingState_* (CellFunctionConstants.h:26-35); 0 = Ready */
    uint8_t cellFunction; /* carried through; the physics path treats every cell as CellFunction_None */
    uint8_t ancestorMutationId;
    uint8_t selected;
    uint32_t creatureId;
    uint32_t mutationId;
    float genomeComplexity;
    uint32_t activationTime;
} alien_cell_to;

/* ParticleTO (TOs.cuh:10-19). */
typedef struct alien_particle_to {
    uint64_t id;
    float energy;
    float pos[2];
    float vel[2];
    uint8_t color;
    uint8_t selected;
    uint8_t pad_[2];
} alien_particle_to;

/*
 * Physics subset of SimulationParameters (source/EngineInterface/SimulationParameters.h:18-206,
 * Motion.h:7-43, SimulationParametersZoneValues.h:10-24).  alien_params_default() fills ALIEN's
 * defaults.  The processes of the reference that this path does not implement are fixed at these
 * values (no field here): radiation sources none (numRadiationSources 0), external energy off,
 * particle transformation off, numZones 0, rigidity 0, cellDeathConsequences None,
 * cellDeathProbability 0.  Cell radiation (CellProcessor::radiation, CellProcessor.cuh:725-773)
 * and particle splitting (RadiationProcessor::splitting, RadiationProcessor.cuh:172-204) are
 * simulated: their fields close the struct.
 */
typedef struct alien_params {
    float timestepSize;                /* SimulationParameters.h:61, default 1.0           */
    int32_t motionType;                /* SimulationParameters.h:62, default Fluid          */
    float smoothingLength;             /* Motion.h:9, default 0.8                          */
    float viscosityStrength;           /* Motion.h:10, default 0.1                         */
    float pressureStrength;            /* Motion.h:11, default 0.1                         */
    float cellMaxCollisionDistance;    /* Motion.h:22, default 1.3                         */
    float cellRepulsionStrength;       /* Motion.h:23, default 0.08                        */
    float friction;                    /* ZoneValues.h:12, default 0.001                   */
    float innerFriction;               /* SimulationParameters.h:65, default 0.3, in [0,1] */
    float cellMaxVelocity;             /* SimulationParameters.h:66, default 2.0           */
    float cellMinDistance;             /* SimulationParameters.h:70, default 0.3           */
    float cellFusionVelocity;          /* ZoneValues.h:23, default 0.6                     */
    float cellMaxBindingEnergy;        /* ZoneValues.h:24, default +inf                    */
    float cellMaxForceDecayProb;       /* SimulationParameters.h:71, default 0.2, in [0,1] */
    float cellMaxBindingDistance[ALIEN_MAX_COLORS]; /* SimulationParameters.h:67, 3.6   */
    float radiationAbsorption[ALIEN_MAX_COLORS];    /* ZoneValues.h:14, 1.0             */
    float cellMaxForce[ALIEN_MAX_COLORS];           /* ZoneValues.h:19, 0.8             */
    float cellMinEnergy[ALIEN_MAX_COLORS];          /* ZoneValues.h:20, 50.0            */
    float radiationProb;               /* SimulationParameters.h:79, default 0.03, in [0,1] */
    float radiationVelocityMultiplier; /* SimulationParameters.h:80, default 1.0           */
    float radiationVelocityPerturbation; /* SimulationParameters.h:81, default 0.5         */
    int32_t radiationMinCellAge[ALIEN_MAX_COLORS];        /* SimulationParameters.h:82, 0  */
    float highRadiationFactor[ALIEN_MAX_COLORS];          /* SimulationParameters.h:85, 0  */
    float highRadiationMinCellEnergy[ALIEN_MAX_COLORS];   /* SimulationParameters.h:86, 500 */
    float radiationCellAgeStrength[ALIEN_MAX_COLORS];     /* ZoneValues.h:17, 0.00002     */
    float particleSplitEnergy[ALIEN_MAX_COLORS];          /* SimulationParameters.h:114, +inf */
} alien_params;

/* World geometry and capacity (GeneralSettings worldSizeX/worldSizeY), and the key of the world's
 * random draws.  The reference draws from a table refilled by the host; here a draw is a function
 * of (randomSeed, timestep, cell id), so a world's run is reproducible from its seed. */
typedef struct alien_world_desc {
    int32_t worldSizeX;
    int32_t worldSizeY;
    uint64_t cellCapacity;     /* 0 = size to the first alien_set_data call     */
    uint64_t particleCapacity; /* 0 = size to the first alien_set_data call     */
    uint64_t randomSeed;
} alien_world_desc;

/* Engine-wide totals (IntegrationTestFramework.cpp:30-40 getEnergy, plus momentum). */
typedef struct alien_totals {
    double cellEnergy;     /* sum of cell energies                  */
    double particleEnergy; /* sum of energy-particle energies       */
    double momentum[2];    /* sum of cell velocities (unit mass)    */
    double kinetic;        /* 1/2 sum |v|^2 over cells              */
    uint64_t numCells;
    uint64_t numParticles;
} alien_totals;

/* Per-colour counts and energies of one world: TimestepStatistics (EngineInterface/
 * RawStatisticsData.h:6-19) restricted to the physics subset (the replicator, virus, colony and
 * genome fields are zero without cell functions, and every cell is a free cell).  Energies are
 * summed in double in a fixed order. */
typedef struct alien_statistics {
    int32_t numCells[7];
    int32_t numParticles[7];
    double totalEnergy[7]; /* cells and energy particles of the colour */
    uint64_t timestep;
} alien_statistics;

typedef struct alien_world alien_world;

/* ---- library ------------------------------------------------------------------------------- */

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* alien_last_error(void);
int alien_abi_version(void);
/* Device name of the current HIP device (SimulationFacade.h:24 getGpuName). */
int alien_get_device_name(char* buf, size_t cap);
/* Fill *p with ALIEN's defaults for the physics subset (see alien_params above). */
void alien_params_default(alien_params* p);

/* ---- one world (the plain path) ------------------------------------------------------------ */

/* _SimulationCudaFacade ctor (SimulationCudaFacade.cuh:37; SimulationCudaFacade.cu:51-87). */
int alien_world_create(const alien_world_desc* desc, const alien_params* params, alien_world** out);
/* ~_SimulationCudaFacade. */
int alien_world_destroy(alien_world* w);
/* setSimulationParameters (SimulationCudaFacade.cuh:74-76); resets fluid density to 1 like
 * prepareForSimulationParametersChanges (SimulationKernelsLauncher.cu:93-97). */
int alien_set_params(alien_world* w, const alien_params* params);
int alien_get_params(const alien_world* w, alien_params* params);
/* setSimulationData (SimulationCudaFacade.cuh:51; SimulationCudaFacade.cu:261-268): replaces all
 * cells and particles.  Positions are wrapped onto the torus, density is set to 1. */
int alien_set_data(alien_world* w, const alien_cell_to* cells, uint64_t numCells,
                   const alien_particle_to* particles, uint64_t numParticles);
/* getSimulationData (SimulationCudaFacade.cuh:46): whole world, cells in ascending id order,
 * bonds as indices into the returned array; particles in ascending id order. */
int alien_get_data(alien_world* w, alien_cell_to* cells, uint64_t cellCap, uint64_t* numCells,
                   alien_particle_to* particles, uint64_t particleCap, uint64_t* numParticles);
int alien_get_counts(alien_world* w, uint64_t* numCells, uint64_t* numParticles);
/* calcTimestep(n) (SimulationCudaFacade.cuh:42; SimulationCudaFacade.cu:128-160). */
int alien_step(alien_world* w, uint64_t steps);
/* get/setCurrentTimestep (SimulationCudaFacade.cuh:86-87). */
uint64_t alien_get_timestep(const alien_world* w);
int alien_set_timestep(alien_world* w, uint64_t timestep);
/* Totals for conservation checks (getRawStatistics subset). */
int alien_get_totals(alien_world* w, alien_totals* out);
/* Per-colour statistics, computed on the device (StatisticsKernels.cu:11-45, cudaUpdateTimestep-
 * Statistics_substep2: cell and particle counts and energies by colour; SimulationFacade.h
 * getStatisticsHistory / getRawStatistics). */
int alien_get_statistics(alien_world* w, alien_statistics* out);
/* alien_get_statistics without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Partial sums and the result live in a device array of the library
 * (about 5.5 MB per GPU, held until the process ends), so the world's cells plus particle slots
 * may number at most 65536 x 256 = 16,777,216 (ALIEN_E_CAPACITY beyond).  Same result as
 * alien_get_statistics, bit for bit.  Calls that overlap on one GPU race on that array. */
int alien_get_statistics_noalloc(alien_world* w, alien_statistics* out);
/* clear (SimulationCudaFacade.cuh:89): removes all cells and particles. */
int alien_clear(alien_world* w);

/* Introspection used by parity tests.  The cell map after the last bin build: for every cell of
 * the current internal order, its spatial bin (row-major over the bin grid) and the cell's id;
 * binSize/binsX/binsY describe the grid. */
int alien_debug_get_map(alien_world* w, int32_t* cellBin, uint64_t* cellId, uint64_t cap,
                        uint64_t* numCells, int32_t* binSize, int32_t* binsX, int32_t* binsY);
/* Neighbour lists of the current state under the step's predicate, as CSR over cells in
 * ascending id order: offsets[numCells+1], neighbour ids in visiting order. */
int alien_debug_neighbors(alien_world* w, uint64_t* offsets, uint64_t offsetsCap,
                          uint64_t* neighborIds, uint64_t idsCap, uint64_t* numIds);

/* Kernel timing (the analog of ALIEN's debug mode, Macros.cuh:76-80, which synchronizes after
 * every launch): while enabled, steps run without graph capture and every kernel launch is
 * bracketed by HIP events on the engine's stream.  alien_get_kernel_timing returns, per kernel
 * name, the number of launches and their summed event time since the last call, and resets. */
typedef struct alien_kernel_time {
    char name[40];
    uint64_t launches;
    double totalMs;
} alien_kernel_time;
int alien_set_kernel_timing(alien_world* w, int32_t enable);
int alien_get_kernel_timing(alien_world* w, alien_kernel_time* out, int32_t cap, int32_t* count);

/* Neighbour-pass path counts since the world's data was last set (no counterpart in ALIEN; a
 * performance diagnostic, results never depend on the path): chunks of bin rows too crowded for
 * the LDS staging (a row above 128 entries or three consecutive rows above 256), which walk their
 * stencils in global memory, and cells with more than 64 candidates, which walk theirs alone. */
int alien_get_path_counts(alien_world* w, uint64_t* fallbackChunks, uint64_t* directCells);

/* ---- batched worlds ------------------------------------------------------------------------ */

typedef struct alien_batch alien_batch;

/* A batch advances many independent worlds of uneven size with one captured graph per call.
 * Each world keeps its own parameters, timestep and results; a world's results do not depend on
 * which other worlds share the batch. */
int alien_batch_create(alien_batch** out);
int alien_batch_destroy(alien_batch* b);
int alien_batch_add_world(alien_batch* b, const alien_world_desc* desc, const alien_params* params,
                          const alien_cell_to* cells, uint64_t numCells,
                          const alien_particle_to* particles, uint64_t numParticles,
                          int32_t* worldIndex);
int alien_batch_num_worlds(const alien_batch* b);
/* Advance every world by `steps` timesteps (one graph launch when graphs are enabled). */
int alien_batch_step(alien_batch* b, uint64_t steps);
int alien_batch_get_data(alien_batch* b, int32_t world, alien_cell_to* cells, uint64_t cellCap,
                         uint64_t* numCells, alien_particle_to* particles, uint64_t particleCap,
                         uint64_t* numParticles);
int alien_batch_get_counts(alien_batch* b, int32_t world, uint64_t* numCells, uint64_t* numParticles);
uint64_t alien_batch_get_timestep(const alien_batch* b, int32_t world);
/* Launch shape: threads per workgroup for the cell kernels: 64, 128, 192 or 256 (results do not
 * depend on it). */
int alien_batch_set_block_size(alien_batch* b, int32_t threads);
int alien_batch_set_use_graph(alien_batch* b, int32_t enable);
/* Autotuner (no counterpart in ALIEN): times trialSteps steps of a device copy of the batch for
 * candidate launch shapes (cell-kernel block size 64/128/256 x neighbour-pass strip cap 64/32/16
 * bins, then bin rows per neighbour-pass chunk 2/4/8 against the default rule, reported as 0)
 * and keeps the fastest, also for worlds added later.  The batch's own worlds do not advance and
 * no result depends on the choice.  Outputs may be NULL. */
int alien_batch_autotune(alien_batch* b, uint32_t trialSteps, int32_t* blockSize, int32_t* stripCap,
                         int32_t* chunkRows, double* msPerStep);
int alien_batch_set_kernel_timing(alien_batch* b, int32_t enable);
int alien_batch_get_kernel_timing(alien_batch* b, alien_kernel_time* out, int32_t cap, int32_t* count);
/* alien_get_path_counts for all worlds of a batch together (no counterpart in ALIEN). */
int alien_batch_get_path_counts(alien_batch* b, uint64_t* fallbackChunks, uint64_t* directCells);

/* ---- serving ------------------------------------------------------------------------------- */

typedef struct alien_server alien_server;

#define ALIEN_JOB_QUEUED 0
#define ALIEN_JOB_RUNNING 1
#define ALIEN_JOB_DONE 2
#define ALIEN_JOB_FAILED 3

typedef struct alien_job_info {
    int32_t state;
    uint64_t stepsDone;
    uint64_t stepsRequested;
    uint64_t numCells;
    uint64_t numParticles;
    double submitTime;   /* seconds, server clock */
    double startTime;    /* admission to the device */
    double finishTime;   /* snapshot ready on the host */
    uint64_t deviceBytes;
} alien_job_info;

/* Continuous-batching server: one controller thread keeps all admitted worlds resident on the
 * device, advances them together in chunks of `chunkSteps`, admits queued jobs whenever their
 * estimated device bytes fit `budgetBytes`, and copies a snapshot out when a job completes.
 * `budgetBytes` bounds the peak: admitting rebuilds the resident set into fresh memory while the
 * old set still holds its own (see alien_server_plan_admission).  A failed rebuild fails only the
 * jobs it was admitting (the resident jobs go on); a failed step fails the jobs of that set. */
int alien_server_create(uint64_t budgetBytes, uint32_t chunkSteps, alien_server** out);
int alien_server_destroy(alien_server* s);
int alien_server_submit(alien_server* s, const alien_world_desc* desc, const alien_params* params,
                        const alien_cell_to* cells, uint64_t numCells,
                        const alien_particle_to* particles, uint64_t numParticles, uint64_t steps,
                        uint64_t* jobId);
int alien_server_poll(alien_server* s, uint64_t jobId, alien_job_info* info);
/* Blocks until the job is done (or failed) or timeoutSeconds elapse (ALIEN_E_STATE on timeout). */
int alien_server_wait(alien_server* s, uint64_t jobId, double timeoutSeconds);
int alien_server_get_snapshot(alien_server* s, uint64_t jobId, alien_cell_to* cells, uint64_t cellCap,
                              uint64_t* numCells, alien_particle_to* particles, uint64_t particleCap,
                              uint64_t* numParticles);
/* Drops a finished job's snapshot, or cancels a queued/running job. */
int alien_server_release(alien_server* s, uint64_t jobId);
/* Estimated device bytes a job of this size needs (the admission figure). */
uint64_t alien_server_estimate_bytes(uint64_t numCells, uint64_t numParticles, int32_t worldSizeX,
                                     int32_t worldSizeY);
/* The admission rule of the server (host arithmetic only): how many of the queued jobs, in order,
 * are admitted now.  oldSetBytes: device bytes of the current resident set (0 if none), which stay
 * allocated while the new set is built; keptBytes: estimated bytes of the resident jobs that go on;
 * queuedBytes[numQueued]: estimates of the queued jobs.  A job is admitted while
 * oldSetBytes + keptBytes + (bytes admitted so far) + its bytes <= budgetBytes; the first job that
 * does not fit stops admission (FIFO). */
uint64_t alien_server_plan_admission(uint64_t budgetBytes, uint64_t oldSetBytes, uint64_t keptBytes,
                                     const uint64_t* queuedBytes, uint64_t numQueued);

/* ---- settings files ---------------------------------------------------------------------- */

/* Processes that a .settings.json may switch on and that this path does not simulate (bits of
 * the `unsimulated` output of alien_settings_read). */
#define ALIEN_SETTINGS_ZONES 0x1u                   /* spots (numZones > 0)                       */
#define ALIEN_SETTINGS_RADIATION_SOURCES 0x2u       /* particle sources                           */
#define ALIEN_SETTINGS_RADIATION 0x4u               /* radiation absorption penalties             */
#define ALIEN_SETTINGS_CELL_DEATH 0x8u              /* death consequences / death probability     */
#define ALIEN_SETTINGS_PARTICLE_SPLIT 0x10u         /* (not reported: particle splitting is simulated) */
#define ALIEN_SETTINGS_RIGIDITY 0x20u               /* rigidity != 0                              */
#define ALIEN_SETTINGS_PARTICLE_TRANSFORMATION 0x40u
#define ALIEN_SETTINGS_ABSORPTION_CONTROL 0x80u     /* features.additional absorption control     */

/* Reads ALIEN's .settings.json (AuxiliaryDataParserService::decodeAuxiliaryData,
 * PersisterInterface/AuxiliaryDataParserService.cpp:1100-1137; the file is written by
 * SerializerService::serializeSimulationToFiles, SerializerService.cpp:850-916): world size and
 * timestep of the general part, the physics subset of the simulation parameters (absent entries
 * take ALIEN's defaults, like ParameterParser::encodeDecode), and in *unsimulated the processes
 * the file switches on that this path does not simulate.  desc->cellCapacity/particleCapacity are
 * 0 and desc->randomSeed 0.  timestep and unsimulated may be NULL. */
int alien_settings_read(const char* json, size_t length, alien_world_desc* desc, alien_params* params,
                        uint64_t* timestep, uint32_t* unsimulated);
/* Writes the same subset in ALIEN's layout (nested objects, every value a string), with the
 * processes this path keeps off written as off; *length is the text length without the final 0
 * (ALIEN_E_CAPACITY, with *length set, if it does not fit `capacity`). */
int alien_settings_write(const alien_world_desc* desc, const alien_params* params, uint64_t timestep, char* buf,
                         size_t capacity, size_t* length);

#ifdef __cplusplus
}
#endif

#endif /* ALIEN_HIP_H */
