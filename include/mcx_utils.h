/*
 * mcx_utils.h - simulation configuration (Config) and the configuration API.
 *
 * Drop-in replacement for the reference's configuration layer
 * (src/mcx_utils.h:165-340).  Front-ends (a CLI, a Python or MATLAB binding)
 * fill a Config through these functions and hand it to mcx_run_simulation()
 * (include/mcx_host.h).  Field names, meanings, defaults and units follow the
 * reference so front-end code written against it compiles unchanged; the
 * OpenCL-only fields (clsource, kernelfile, compileropt, optlevel) are kept
 * for source compatibility and ignored.
 *
 * Error convention (reference src/mcx_utils.c:1298-1305): mcx_error() prints
 * the message and exits in a standalone program.  A caller that must not exit
 * (a language binding) arms a recovery point with the mcx_try_* entry points,
 * which return a non-zero code and leave the message in mcx_last_error().
 */
#ifndef MCX_AMD_UTILS_H
#define MCX_AMD_UTILS_H

#include <stddef.h>
#include <stdio.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 16-byte aligned host vector types, layout-compatible with HIP's float4/uint4 */
typedef struct MCXFloat4 {
    float x, y, z, w;
} __attribute__((aligned(16))) MCXFloat4;

typedef struct MCXUInt4 {
    unsigned int x, y, z, w;
} __attribute__((aligned(16))) MCXUInt4;

typedef struct MCXFloat3 {
    float x, y, z;
} MCXFloat3;

typedef struct MCXUInt3 {
    unsigned int x, y, z;
} MCXUInt3;

#if !defined(__HIPCC__)
typedef MCXFloat4 float4;
typedef MCXUInt4 uint4;
typedef MCXFloat3 float3;
typedef MCXUInt3 uint3;
#endif

#define MCX_VERSION        "v2025.10-mi355x"
#define MAX_FULL_PATH      2048
#define MAX_PATH_LENGTH    1024
#define MAX_SESSION_LENGTH 64
#define MAX_DEVICE         256
#define MAX_PROP           2000
#define MAX_DETECTORS      2000

#define MCX_ERROR(id, msg) mcx_error(id, msg, __FILE__, __LINE__)
#define MCX_ASSERT(x)      mcx_assess((x), "assert error", __FILE__, __LINE__)

/* constants shared with the reference (src/mcx_const.h:29-118) */
#define ONE_PI             3.1415926535897932f
#define TWO_PI             6.28318530717959f
#define EPS                1e-10f
#define R_C0               3.335640951981520e-12f   /* 1/c0 in s/mm */
#define VERY_BIG           1e10f
#define SEED_FROM_FILE     -999
#define FILL_MAXDETPHOTON  3
#define SIGN_BIT           0x80000000U
#define OUTSIDE_VOLUME_MIN 0xFFFFFFFFu
#define OUTSIDE_VOLUME_MAX 0x7FFFFFFFu
#define DET_MASK           0x80000000u
#define MED_MASK           0x7FFFFFFFu
#define NO_LAUNCH          9999

#define MEDIA_ASGN_F2H     96
#define MEDIA_2LABEL_SPLIT 97
#define MCX_NANGLES        181  /* scattering angles of a polarised medium's Mie table, 0..180 degrees */
#define MEDIA_2LABEL_MIX   98
#define MEDIA_LABEL_HALF   99
#define MEDIA_AS_F2H       100
#define MEDIA_MUA_FLOAT    101
#define MEDIA_AS_HALF      102
#define MEDIA_ASGN_BYTE    103
#define MEDIA_AS_SHORT     104
/* extension: per-voxel float32 {mua, mus} pairs (two uint32 words per voxel),
   label-free, used by the differentiable replay path */
#define MEDIA_MUA_MUS_FLOAT 200

#define MCX_DEBUG_RNG       1
#define MCX_DEBUG_MOVE      2
#define MCX_DEBUG_PROGRESS  4
#define MCX_DEBUG_MOVE_ONLY 8
#define MCX_DEBUG_REC_LEN   6

#define MCX_SRC_PENCIL      0
#define MCX_SRC_ISOTROPIC   1
#define MCX_SRC_CONE        2
#define MCX_SRC_GAUSSIAN    3
#define MCX_SRC_PLANAR      4
#define MCX_SRC_PATTERN     5
#define MCX_SRC_FOURIER     6
#define MCX_SRC_ARCSINE     7
#define MCX_SRC_DISK        8
#define MCX_SRC_FOURIERX    9
#define MCX_SRC_FOURIERX2D  10
#define MCX_SRC_ZGAUSSIAN   11
#define MCX_SRC_LINE        12
#define MCX_SRC_SLIT        13
#define MCX_SRC_PENCILARRAY 14
#define MCX_SRC_PATTERN3D   15
#define MCX_SRC_HYPERBOLOID_GAUSSIAN 16
#define MCX_SRC_RING        17

#define SAVE_DETID(a)   ((a) & 0x1)
#define SAVE_NSCAT(a)   ((a) >> 1 & 0x1)
#define SAVE_PPATH(a)   ((a) >> 2 & 0x1)
#define SAVE_MOM(a)     ((a) >> 3 & 0x1)
#define SAVE_PEXIT(a)   ((a) >> 4 & 0x1)
#define SAVE_VEXIT(a)   ((a) >> 5 & 0x1)
#define SAVE_W0(a)      ((a) >> 6 & 0x1)
#define SAVE_IQUV(a)    ((a) >> 7 & 0x1)

enum TOutputType {otFlux, otFluence, otEnergy, otJacobian, otWP, otDCS, otRF, otL, otRFmus, otWLTOF, otWPTOF,
                  otAdjoint, otAdjointDcoeff, otAdjointMus, otAdjointMusp, otAdjointMuaD, otAdjointMuaMusp
                 };
#define MCX_IS_ADJOINT_TYPE(t)      ((int)(t) >= (int)otAdjoint)
#define MCX_IS_DUAL_ADJOINT_TYPE(t) ((int)(t) >= (int)otAdjointMuaD)
enum TMCXParent {mpStandalone, mpMATLAB, mpPython};
enum TOutputFormat {ofMC2, ofNifti, ofAnalyze, ofUBJSON, ofTX3, ofJNifti, ofBJNifti};
enum TDeviceVendor {dvUnknown, dvNVIDIA, dvAMD, dvIntel, dvIntelGPU, dvAppleCPU, dvAppleGPU, dvArmGPU};
enum TBoundary {bcUnknown, bcReflect, bcAbsorb, bcMirror, bcCyclic};

typedef struct MCXMedium {
    float mua;      /* absorption coefficient, 1/mm */
    float mus;      /* scattering coefficient, 1/mm */
    float g;        /* anisotropy */
    float n;        /* refractive index */
} __attribute__((aligned(16))) Medium;

typedef struct MCXPolarizeMedium {
    float mua, r, rho, nsph, nmed;
} POLMedium;

typedef struct MCXExtraSource {
    MCXFloat4 srcpos, srcdir, srcparam1, srcparam2;
} ExtraSrc;

/* 256-byte .mch header (reference src/mcx_utils.h:113-130) */
typedef struct MCXHistoryHeader {
    char magic[4];
    unsigned int version;
    unsigned int maxmedia;
    unsigned int detnum;
    unsigned int colcount;
    unsigned int totalphoton;
    unsigned int detected;
    unsigned int savedphoton;
    float unitinmm;
    unsigned int seedbyte;
    float normalizer;
    int respin;
    unsigned int srcnum;
    unsigned int savedetflag;
    unsigned int totalsource;
    int reserved[1];
} History;

typedef struct PhotonReplay {
    int* detid;
    void* seed;
    float* weight;
    float* tof;
} Replay;

typedef struct MCXGPUInfo {
    char name[MAX_SESSION_LENGTH];
    int id;
    int devcount;
    int platformid;
    int major;
    int minor;
    size_t globalmem;
    size_t constmem;
    size_t sharedmem;
    int regcount;
    int clock;
    int sm;                 /* compute units */
    int core;               /* stream processors (CU x 64) */
    size_t autoblock;
    size_t autothread;
    unsigned int maxgate;
    int maxmpthread;
    int iscpu;
    enum TDeviceVendor vendor;
} GPUInfo;

typedef struct MCXConfig {
    size_t nphoton;
    unsigned int nblocksize;
    unsigned int nthread;
    int seed;

    MCXFloat4 srcpos;
    MCXFloat4 srcdir;
    MCXFloat4 srciquv;
    float tstart;
    float tstep;
    float tend;
    MCXFloat4 steps;

    MCXUInt4 dim;
    MCXUInt4 crop0;
    MCXUInt4 crop1;
    unsigned int medianum;
    unsigned int polmedianum;
    unsigned int detnum;
    unsigned int maxdetphoton;
    float detradius;
    float sradius;

    Medium* prop;
    POLMedium* polprop;
    MCXFloat4* detpos;
    MCXFloat4* detdir;
    MCXFloat4* smatrix;

    unsigned int maxgate;
    int respin;
    unsigned int printnum;
    int gpuid;

    unsigned int* vol;
    char session[MAX_SESSION_LENGTH];
    char isrowmajor;
    char isreflect;
    char isref3;
    char isrefint;
    char isnormalized;
    char issavedet;
    char issave2pt;
    char isgpuinfo;
    char isspecular;
    char isverbose;
    char issrcfrom0;
    char ismomentum;
    char internalsrc;
    char isdumpmask;
    char issaveseed;
    char istrajstokes;
    char issaveexit;
    char isatomic;
    char issaveref;
    char isdumpjson;
    int zipid;
    char srctype;
    char autopilot;
    char outputtype;
    char outputformat;
    float minenergy;
    float unitinmm;
    float omega;
    float lambda;
    FILE* flog;
    History his;
    double energytot;
    double energyabs;
    double energyesc;
    char rootpath[MAX_PATH_LENGTH];
    char kernelfile[MAX_SESSION_LENGTH];
    char compileropt[MAX_PATH_LENGTH];
    char* shapedata;
    char* extrajson;
    char* clsource;
    int maxvoidstep;
    int voidtime;
    MCXFloat4 srcparam1;
    MCXFloat4 srcparam2;
    unsigned int srcnum;
    float* srcpattern;
    char deviceid[MAX_DEVICE];
    float workload[MAX_DEVICE];
    float* exportfield;
    float* exportjacob;
    float* exportdetected;
    unsigned int debuglevel;
    char faststep;
    float normalizer;
    Replay replay;
    void* seeddata;
    int replaydet;
    char seedfile[MAX_PATH_LENGTH];
    char jsonfile[MAX_PATH_LENGTH];
    unsigned int maxjumpdebug;
    unsigned int debugdatalen;
    unsigned int gscatter;
    float* exportdebugdata;
    unsigned int savedetflag;
    unsigned int detectedcount;
    unsigned int runtime;
    int parentid;
    unsigned int optlevel;
    unsigned int mediabyte;
    char bc[13];
    unsigned int nphase;
    float* invcdf;
    unsigned int nangle;
    float* angleinvcdf;
    int srcid;
    unsigned int extrasrclen;
    ExtraSrc* srcdata;
    /* ---- MI355X additions (no reference counterpart) ---- */
    int mathmode;            /* 0: fast hardware transcendentals, 1: parity (bit-reproducible) */
    float kernelms;          /* kernel-only time of the last run, from HIP events (ms) */
    float jobms;             /* transport kernel plus fluence finalisation, from HIP events (ms) */
    double runtimems;        /* wall-clock of the photon loop of the last run (ms) */
    double* exportfield64;   /* optional: unnormalised double-precision field (fieldlen entries, caller-owned) */
    float* muamus;           /* optional: per-voxel {mua, mus} in 1/mm (2 x nvox floats, x fastest, caller-owned);
                                voxels with a non-zero label take mua and mus from here, g and n from the label */
    float* exportdetweight;  /* optional: exit weight of each detected photon (maxdetphoton floats, caller-owned) */
    float* replayinweight;   /* optional: replay weight of each row of detphotons, overriding the weight computed
                                from partial path lengths (needed with muamus) */
    float* exportfield32;    /* optional: the unnormalised field converted to float32 on the device (fieldlen
                                entries, caller-owned, overwritten); with exportfield NULL no normalised copy
                                is made, so large replay grids are read back once at half the bytes */
} Config;

/* ---- configuration API; each line names the reference declaration it replaces ---- */
void mcx_initcfg(Config* cfg);                                              /* src/mcx_utils.h:295 */
void mcx_clearcfg(Config* cfg);                                             /* src/mcx_utils.h:296 */
void mcx_error(const int id, const char* msg, const char* file, const int linenum); /* src/mcx_utils.h:288 */
void mcx_assess(const int id, const char* msg, const char* file, const int linenum); /* src/mcx_utils.h:289 */
void mcx_readconfig(char* fname, Config* cfg);                              /* src/mcx_utils.h:293 */
void mcx_loadconfig(FILE* in, Config* cfg);                                 /* src/mcx_utils.h:291 */
/* replaces int mcx_loadjson(cJSON* root, Config* cfg) (src/mcx_utils.h:313): the
   parsed-tree type is internal here, so the entry point takes the JSON text */
int  mcx_loadjson_string(const char* json, Config* cfg);
void mcx_loadbenchmark(char* key, Config* cfg);                             /* src/mcx_utils.h:335 */
void mcx_prepdomain(char* filename, Config* cfg);                           /* src/mcx_utils.h:305 */
void mcx_preprocess(Config* cfg);                                           /* src/mcx_utils.c:1521 (not in the reference header) */
void mcx_prep_polarized(Config* cfg);                                       /* src/mcx_utils.h:340 */
/* replaces void Mie(double x, double mx, const double* mu, float4* smatrix, double* qsca,
   double* g) (src/mcx_mie.h:42): nangle angles, smatrix as 4 floats per angle, an error
   code instead of an exit for x outside (0, 20000] */
int  mcx_mie(double x, double m, int nangle, const double* mu, float* smatrix, double* qsca, double* g);
void mcx_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte); /* src/mcx_utils.h:337 */
void mcx_append_detector_sources(Config* cfg);                              /* the detector-as-source append of
                                                                               src/mcx_utils.c:1902-1925 and
                                                                               src/pmcxcl.cpp:1179-1198 (inline there),
                                                                               for bindings after mcx_validatecfg */
void mcx_replayinit(Config* cfg, float* detps, int dimdetps[2], int seedbyte); /* src/mcx_utils.h:336 */
void mcx_maskdet(Config* cfg);                                              /* src/mcx_utils.h:304 */
void mcx_normalize(float field[], float scale, int fieldlen, int option, int pidx, int srcnum); /* src/mcx_utils.h:300 */
void mcx_kahanSum(float* sum, float* kahanc, float input);                  /* src/mcx_utils.h:326 */
float mcx_updatemua(unsigned int mediaid, Config* cfg);                     /* src/mcx_utils.h:327 */
int  mcx_pack_media(unsigned int format, const float* in, size_t nvox, float unitinmm, unsigned int* out);
                                                                            /* the packing steps of src/mcx_utils.c:3597-3681
                                                                               and src/pmcxcl.cpp:243-389 (inline there) */
void mcx_createfluence(float** fluence, Config* cfg);                       /* src/mcx_utils.h:306 */
void mcx_clearfluence(float** fluence);                                     /* src/mcx_utils.h:307 */
void mcx_convertrow2col(unsigned int** vol, MCXUInt4* dim);                 /* src/mcx_utils.h:309 */
void mcx_cleargpuinfo(GPUInfo** gpuinfo);                                   /* src/mcx_utils.h:290 */
void mcx_flush(Config* cfg);                                                /* src/mcx_utils.h:324 */
int  mcx_lookupindex(char* key, const char* index);                         /* src/mcx_utils.h:317 */
int  mcx_keylookup(char* key, const char* table[]);                         /* src/mcx_utils.h:314 */
int  mcx_parsedebugopt(char* debugopt, const char* debugflag);              /* src/mcx_utils.h:318 */
void mcx_savedata(float* dat, size_t len, Config* cfg);                     /* src/mcx_utils.h:286 */
void mcx_savedetphoton(float* ppath, void* seeds, int count, int doappend, Config* cfg); /* src/mcx_utils.h:312 */
void mcx_savenii(float* dat, size_t len, char* name, int type32bit, int outputformatid, Config* cfg); /* src/mcx_utils.h:287 */
void mcx_savejnii(float* vol, int ndim, unsigned int* dims, float* voxelsize, char* name, int isfloat, int iscol,
                  Config* cfg);                                             /* src/mcx_utils.h:330 */
void mcx_savebnii(float* vol, int ndim, unsigned int* dims, float* voxelsize, char* name, int isfloat, int iscol,
                  Config* cfg);                                             /* src/mcx_utils.h:331 */
void mcx_savejdet(float* ppath, void* seeds, unsigned int count, int doappend, Config* cfg); /* src/mcx_utils.h:332 */
void mcx_replayprep(int* detid, float* ppath, History* his, Config* cfg);   /* src/mcx_utils.h:333 */
void mcx_loadseedfile(Config* cfg);                                         /* src/mcx_utils.h:325 */
void mcx_loadseedjdat(char* filename, Config* cfg);                         /* src/mcx_utils.c:3707 (not in the reference header) */
/* replaces int mcx_parse_shapestring(Grid3D* g, char* shapedata) (src/mcx_shapes.h:56):
   rasterises the shapes into cfg->vol/cfg->dim directly */
int  mcx_parse_shapestring(Config* cfg, const char* shapedata);
void mcx_parsecmd(int argc, char* argv[], Config* cfg);                     /* src/mcx_utils.h:297 */
void mcx_usage(Config* cfg, char* exename);                                 /* src/mcx_utils.h:298 */

/* record geometry of a detected photon (reference src/mcx_host.cpp:494-496) */
unsigned int mcx_detreclen(const Config* cfg);
unsigned int mcx_partialdata(const Config* cfg);

/* non-exiting variants for language bindings: return 0 on success, the
   mcx_error id otherwise; the message is kept in mcx_last_error() */
int mcx_try_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte);
int mcx_try_loadjson_string(const char* json, Config* cfg);
int mcx_try_loadbenchmark(const char* key, Config* cfg);
int mcx_try_readconfig(const char* fname, Config* cfg);
int mcx_try_parse_shapestring(Config* cfg, const char* shapedata);
int mcx_try_parsecmd(int argc, char* argv[], Config* cfg);
int mcx_try_savedata(float* dat, size_t len, Config* cfg);
int mcx_try_savedetphoton(float* ppath, void* seeds, int count, int doappend, Config* cfg);
const char* mcx_last_error(void);

/* ABI self-description, used by bindings to verify their Config mirror */
size_t mcx_config_sizeof(void);
long   mcx_config_offsetof(const char* field);
const char* mcx_version_string(void);

/* lists of names (reference src/mcx_utils.c:132-199) */
extern const char* srctypeid[];
extern const char outputtype[];
extern const char boundarycond[];
extern const char boundarydetflag[];
extern const char saveflag[];
extern const char debugflag[];
extern const char* benchname[];

#ifdef __cplusplus
}
#endif

#endif
