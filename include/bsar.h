/*
 * bsar.h -- C ABI of the MI355X-native bistatic-SAR imaging hot path.
 *
 * The reference (DavidAtenciaMondragon/c-_bistatic_sar) exposes this path as
 * C++ free functions and an inline OpenMP loop; each entry point below names
 * the reference interface it replaces (file:line).  Plain pointers and sizes
 * only: no C++ or torch types cross this boundary.
 *
 * Conventions
 *   - complex arrays are interleaved (re, im) float64, as in the reference's
 *     std::complex<double> vectors and its binary files;
 *   - every call returns BSAR_OK (0) or a negative bsar_status; the message
 *     of the last failure on a context is bsar_last_error(ctx);
 *   - a context owns one HIP stream on one GPU and is not thread-safe: use
 *     one context per host thread (the reference runs these functions from
 *     OpenMP threads, proc_test_script.cpp:451);
 *   - host buffers are owned by the caller; device memory is owned by the
 *     context or the dataset that allocated it.
 */
#ifndef BSAR_H
#define BSAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSAR_ABI_VERSION 1

#if defined(__GNUC__)
#define BSAR_API __attribute__((visibility("default")))
#else
#define BSAR_API
#endif

typedef enum {
    BSAR_OK = 0,
    BSAR_ERR_INVALID = -1,   /* bad argument, shape or size */
    BSAR_ERR_DEVICE = -2,    /* HIP runtime error or no device */
    BSAR_ERR_NOMEM = -3,     /* device or host allocation failed */
    BSAR_ERR_IO = -4,        /* file could not be opened, read or written */
    BSAR_ERR_FORMAT = -5,    /* malformed binary/CSV file */
    BSAR_ERR_CAPACITY = -6,  /* serving: job does not fit the HBM cap */
    BSAR_ERR_STATE = -7      /* call out of order (e.g. no range data loaded) */
} bsar_status;

typedef enum {
    BSAR_FP64 = 0,   /* plain form: every step in FP64, reference operation order (R1 sums: squared
                        distances with two fused terms, square roots within 0.51 ulp) */
    BSAR_MIXED = 1   /* FP64 refraction solve + delay/phase, FP32 interpolation,
                        Kahan-compensated FP32 accumulation, tabulated R1 means */
} bsar_precision;

/* Radar/system/environment constants that the reference fixes at compile
 * time (radar_init.cpp:5-89) or reads from environment.txt
 * (raw_data_processing.cpp:14-46).  Runtime values here. */
typedef struct {
    double fc;       /* carrier [Hz]            RadarTx.FreqPortadora, radar_init.cpp:9 */
    double c;        /* propagation speed [m/s] System.VelocidadeLuz, radar_init.cpp:74 */
    double fs;       /* receiver ADC rate [Hz]  RadarRx.fs, radar_init.cpp:44 (before upsampling) */
    double n1, n2;   /* refractive indices      Environment, raw_data_processing.cpp:18-19 */
    int upsample;    /* fast-time factor        factorInterp, proc_test_script.cpp:220 */
    int reserved;
} bsar_params;

/* DEM as ProcessedDEM holds it (radar_params.h:75-79; loaded at
 * proc_test_script.cpp:276-278): X_vec[nx], Y_vec[ny], Z_DEM[zrows][zcols]
 * with Z[j][i] the height at (Y_vec[j], X_vec[i]).  Entries beyond
 * zrows/zcols read as 0.0 (raw_data_processing.cpp:649-652). */
typedef struct {
    const double* x; int nx;
    const double* y; int ny;
    const double* z; int zrows; int zcols;
} bsar_dem;

typedef struct bsar_ctx bsar_ctx;
typedef struct bsar_dataset bsar_dataset;

/* Filled by bsar_backproject when non-NULL. */
typedef struct {
    double seconds;        /* device time of the back-projection launches (HIP events) */
    double voxel_pulses;   /* voxels x pulses back-projected */
    int64_t voxels;
    int32_t pulses;
    int32_t launches;      /* kernel launches used */
    double r1_exact_frac;  /* MIXED: fraction of R1 means computed exactly (outside the table) */
} bsar_report;

/* Transmitted pulse of createCHIRP (raw_data_processing.cpp:257-290):
 * RadarTx.FreqMenor, FreqMaior, DuracaoPulso, fs (radar_init.cpp:17-27). */
typedef struct { double f0, f1, T, fs; } bsar_chirp_params;

/* ---- context ------------------------------------------------------------ */
BSAR_API int  bsar_ctx_create(int device, bsar_ctx** out);
BSAR_API void bsar_ctx_destroy(bsar_ctx* ctx);
BSAR_API const char* bsar_last_error(const bsar_ctx* ctx);
BSAR_API const char* bsar_status_string(int status);
BSAR_API int  bsar_abi_version(void);
BSAR_API int  bsar_ctx_synchronize(bsar_ctx* ctx);

/* ---- stand-alone kernels ------------------------------------------------ */

/* interpolateDEMElevation, raw_data_processing.cpp:615-680 (declared
 * raw_data_processing.h:146), batched: z[k] = DEM(xy[2k], xy[2k+1]). */
BSAR_API int bsar_dem_interp(bsar_ctx* ctx, const bsar_dem* dem, const double* xy, int64_t m, double* z);

/* calculateRefractionPoint, raw_data_processing.cpp:455-572 (declared
 * raw_data_processing.h:135-139), batched over m independent legs
 * legs[k] = (x1,y1,z1,x2,y2,z2); out[k] = refraction point (x,y,z).  The
 * argmin is bit-identical to the reference's FP64 grid search. */
BSAR_API int bsar_refraction_points(bsar_ctx* ctx, const bsar_dem* dem, const double* legs, int64_t m,
                           double n1, double n2, double* out);

/* The same refraction points (calculateRefractionPoint, raw_data_processing.cpp:
 * 455-572), computed without screening: one wavefront per leg, its lanes
 * over every candidate of each level, reduced to the argmin with the lowest
 * reference index among ties.  Bit-identical to bsar_refraction_points;
 * *seconds (may be null) receives the kernel's device time.  m < 2^33. */
BSAR_API int bsar_refraction_points_wave(bsar_ctx* ctx, const bsar_dem* dem, const double* legs, int64_t m,
                                         double n1, double n2, double* out, double* seconds);

/* Refraction points of calculaSlantRangeFermat, raw_data_processing.cpp:112-
 * 255, for V targets and the dataset's N pulses:
 * ida[v][i] = refraction(Tx_i -> P_v, n1, n2), volta[v][i] = refraction(P_v -> Rx_i, n2, n1).
 * pts is [V][3]; ida and volta are [V][N][3]. */
BSAR_API int bsar_slant_range_fermat(bsar_ctx* ctx, const bsar_dataset* ds, const double* pts, int64_t V,
                            double* ida, double* volta);

/* processRawData, raw_data_processing.cpp:292-452 (RawDataResult,
 * raw_data_processing.h:38-43), on the GPU over the dataset's DEM, antennas
 * and constants (fc, c, n1, n2; params.fs is RadarRx.fs): for each of ns
 * point scatterers scat[ns][3] with complex amplitude rcs[ns] and each pulse,
 * the refraction points and slant ranges of calculaSlantRangeFermat, the
 * response rcs * exp(-i 2 pi / lambda * L) at row 1 + round(L / c * fs)
 * (:322-361; rows outside [0, S) dropped; the reference simulates one Target,
 * several scatterers add in order); raw = fftConvolution of each pulse with
 * the fftshifted chirp reference (:371-421) and root = fftCorrelation of raw
 * with it (pulse compression, :431-446), FP64.  S (System.IndiceMaximo) is a
 * power of two at least the chirp length.  Outputs (any may be NULL) are
 * [N][S] complex, pulse-major (RawDataResult's column i is row i here: the
 * range-fastest layout rootData_complex.bin and the back-projection use). */
BSAR_API int bsar_process_raw(bsar_ctx* ctx, const bsar_dataset* ds, const double* scat, const double* rcs, int ns,
                     const bsar_chirp_params* chirp, int S, double* aux, double* raw, double* root);

/* Same with caller-owned device buffers (d_pts [V][3], d_ida / d_volta
 * [V][N][3]) for callers that keep the refraction points on the GPU; rep
 * (optional) receives the kernel time (voxel_pulses = V*N, two solves each). */
BSAR_API int bsar_slant_range_fermat_device(bsar_ctx* ctx, const bsar_dataset* ds, const double* d_pts, int64_t V,
                                   double* d_ida, double* d_volta, bsar_report* rep);

/* splineInterpolation, interpolation.cpp:4-51 (declared interpolation.h:8),
 * over N pulses at once as proc_test_script.cpp:326-350 drives it:
 * in [N][S] complex (range fastest), out [N][S*factor] complex.  S >= 2. */
BSAR_API int bsar_upsample(bsar_ctx* ctx, const double* in, int S, int N, int factor, double* out);

/* FFTProcessor::fftCorrelation (conj = 1, pulse compression) and
 * fftConvolution (conj = 0), fft_utils.cpp:63-137 (declared fft_utils.h:
 * 35-60), applied to N columns of length S against one reference of length
 * S, as processRawData drives them (raw_data_processing.cpp:406-446).
 * sig and out are [N][S] complex (column = pulse, range fastest); S must be
 * a power of two (the reference pads to one, fft_utils.cpp:71).
 * precision: BSAR_FP64 runs complex128, BSAR_MIXED complex64 arithmetic. */
BSAR_API int bsar_range_compress(bsar_ctx* ctx, const double* sig, int S, int N, const double* ref,
                        int conj, int precision, double* out);

/* FFTProcessor::fft (inverse = 0) and FFTProcessor::ifft (inverse = 1),
 * fft_utils.cpp:11-60 (declared fft_utils.h:19-29), in place on N columns
 * of length S: data is [N][S] complex.  Forward: X[k] = sum_n x[n]
 * exp(-2 pi i n k / S); inverse: conj(FFT(conj(x))) / S.  As in the
 * reference, S = 1 or S not a power of two leaves the forward transform a
 * no-op (the reference prints a warning) and the inverse a division by S.
 * The GPU computes each twiddle directly, where the reference multiplies
 * them up by recurrence: results agree to ~3e-17 * S of the peak. */
BSAR_API int bsar_fft(bsar_ctx* ctx, double* data, int S, int N, int inverse);

/* Same filter on a caller-owned device buffer, in place: d_sig is [N][S]
 * complex128 (BSAR_FP64) or complex64 (BSAR_MIXED), S a power of two; ref is
 * the host reference (complex128, length S).  rep (optional) receives the
 * device time of the batched FFT -> multiply -> IFFT over the columns; the
 * reference's own spectrum (its FFT, conjugation and reordering) is made
 * before that interval and is not included. */
BSAR_API int bsar_range_compress_device(bsar_ctx* ctx, void* d_sig, int S, int N, const double* ref, int conj,
                               int precision, bsar_report* rep);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bsar_range_compress_device with every buffer beyond the
 * caller's columns (reference spectrum, FFT buffers, twiddle tables) taken
 * from a fixed 4 MiB scratch array per GPU instead of hipMalloc.  Same
 * kernels, same bits.  Sizes whose buffers exceed the scratch return
 * BSAR_ERR_CAPACITY: up to S = 16384 (complex128) / 32768 (complex64) every N
 * fits (about 60 S bytes); longer columns take N * S elements more.  All calls
 * on one GPU share its scratch: calls that overlap on one GPU (two host
 * threads or contexts) race.  The call returns after its kernels end. */
BSAR_API int bsar_range_compress_device_noalloc(bsar_ctx* ctx, void* d_sig, int S, int N, const double* ref,
                                                int conj, int precision, bsar_report* rep);

/* createCHIRP, raw_data_processing.cpp:257-290: writes L = int(T*fs)+1
 * samples when out != NULL; returns L (> 0) or a negative status. */
BSAR_API int bsar_chirp(double f0, double f1, double T, double fs, double* out);

/* ---- datasets (what proc_test_script.cpp:225-292 loads) ----------------- */


/* DEM, Tx/Rx positions ([n_pulses][3], Tx_positions.csv / Rx_positions.csv)
 * and constants.  The dataset keeps them resident on the context's GPU. */
BSAR_API int bsar_dataset_create(bsar_ctx* ctx, const bsar_params* p, const bsar_dem* dem,
                        const double* tx, const double* rx, int n_pulses, bsar_dataset** out);
BSAR_API void bsar_dataset_destroy(bsar_dataset* ds);

/* rootData_complex.bin payload: n_data pulses of S complex samples, range
 * fastest (flat index r + p*S, proc_test_script.cpp:330).  Uploaded and
 * upsampled by params.upsample on the GPU (proc_test_script.cpp:326-350);
 * the upsampled block stays resident. */
BSAR_API int bsar_dataset_load_rootdata(bsar_ctx* ctx, bsar_dataset* ds, const double* root, int S, int n_data);

/* Same, from a caller-owned device buffer: d_root is [n_data][S] complex128
 * (BSAR_FP64) or complex64 (BSAR_MIXED, widened exactly to FP64 as the
 * spline reads it), e.g. the output of bsar_range_compress_device. */
BSAR_API int bsar_dataset_load_rootdata_device(bsar_ctx* ctx, bsar_dataset* ds, const void* d_root, int S, int n_data,
                                      int precision);

/* Same, when the samples are already upsampled ([n_data][S4] complex). */
BSAR_API int bsar_dataset_load_upsampled(bsar_ctx* ctx, bsar_dataset* ds, const double* rcomp, int S4, int n_data);

/* Device memory the dataset holds, in bytes. */
BSAR_API int64_t bsar_dataset_device_bytes(const bsar_dataset* ds);

/* ---- back-projection (proc_test_script.cpp:395-798) -------------------- */

/* Image over the grid axes gx[nx], gy[ny], gz[nz] (grid_{x,y,z}Axis.csv).
 * out is [nz][ny][nx] complex, x fastest -- the order export3DComplexMatrix
 * writes (data_export.cpp:241-256); the reference's output[x][y][z] is
 * out[(z*ny + y)*nx + x]. */
BSAR_API int bsar_backproject(bsar_ctx* ctx, const bsar_dataset* ds,
                     const double* gx, int nx, const double* gy, int ny, const double* gz, int nz,
                     int precision, double* out, bsar_report* rep);

/* Same, into device memory the caller owns (d_out: nx*ny*nz complex double
 * on the context's device); no host transfer of the image. */
BSAR_API int bsar_backproject_device(bsar_ctx* ctx, const bsar_dataset* ds,
                            const double* gx, int nx, const double* gy, int ny, const double* gz, int nz,
                            int precision, void* d_out, bsar_report* rep);

/* ---- files (stage boundary build/output/proc/) ------------------------- */

/* readBinary<double>, binary_reader.cpp:9-99: header (u32 ndims, u32
 * isComplex, u32 typeLen, char[16], u64 dims[ndims]) + float64 payload.
 * With buf == NULL only the header is parsed.  dims receives up to
 * max_dims sizes; *count the element count; real files read as im = 0. */
BSAR_API int bsar_read_matrix(const char* path, uint64_t* dims, int max_dims, int* ndims,
                     int* is_complex, double* buf, uint64_t* count);

/* export3DComplexMatrix, data_export.cpp:180-309: 3-D complex file (dims
 * nx, ny, nz; data x fastest) plus <base>_grid_{x,y,z}.csv.  img is
 * [nz][ny][nx] complex as bsar_backproject writes it. */
BSAR_API int bsar_write_image(const char* path, const double* img, int nx, int ny, int nz,
                     const double* gx, const double* gy, const double* gz);

/* exportComplexMatrix, data_export.cpp:71-138: 2-D complex file,
 * dims (rows, cols), data row-major. */
BSAR_API int bsar_write_matrix2d(const char* path, const double* data, uint64_t rows, uint64_t cols);

/* ---- scenario generators and DEM files (host only, no context) ----------
 * Context-free calls; on failure bsar_file_error() holds the calling
 * thread's last message. */

BSAR_API const char* bsar_file_error(void);

/* funcao_espiral, funcao_espiral.cpp:5-49: spiral from radius r_major to
 * r_minor (the reference passes RaioMenorEsp second, RaioMaiorEsp third,
 * trajectory_generator.cpp:6-12) over `turns` turns, height h_high -> h_low,
 * tangential speed vt, one position per 1/prf s while t <= T.  mirror_xy
 * negates x and y: generateReceiverTrajectory, trajectory_generator.cpp:17-33.
 * With px/py/pz NULL only *count is returned.  BSAR_ERR_INVALID when the
 * reference's time loop would not end (prf <= 0, zero radial rate with
 * distinct radii) or cap is too small. */
BSAR_API int bsar_spiral(double turns, double r_minor, double r_major, double h_high, double h_low, double vt,
                         double prf, int mirror_xy, double* px, double* py, double* pz, int64_t cap,
                         int64_t* count);

/* One axis of crearGrid, crear_grid.cpp:21-34: center-half, +step, ... while
 * <= center+half+step/2 (a running sum, as the reference builds it).  With
 * out NULL only *count is returned. */
BSAR_API int bsar_grid_axis(double center, double step, double half, double* out, int cap, int* count);

/* A DEM read from a file.  dem's arrays are allocated by the library and
 * released by bsar_dem_grid_free. */
typedef struct {
    bsar_dem dem;
    double lon0, lat0;   /* geographic GeoTIFF: raster centre [deg] (NaN otherwise) */
    double e0, n0;       /* GeoTIFF: the centre (the origin of x, y) as UTM easting/northing
                            (geographic) or model metres (projected) */
    int utm_zone;        /* geographic GeoTIFF: UTM zone of the centre, negative south */
    int nodata_count;    /* GeoTIFF: nodata pixels, replaced by the mean of the valid ones */
    int source;          /* BSAR_DEM_SOURCE_* */
    int reserved;
} bsar_dem_grid;

enum {
    BSAR_DEM_FILE_REFERENCE = 0, /* loadDEM without GDAL, load_dem_simple.cpp:9-224: a name containing
                                    ".tif"/".TIF" or a file that cannot be opened gives the flat 32 x 32
                                    DEM over [-500, 500] m; anything else is read as an ESRI ASCII grid */
    BSAR_DEM_FILE_GEOTIFF = 1    /* decode a single-band GeoTIFF to local metres about its centre */
};
enum { BSAR_DEM_SOURCE_DEFAULT = 0, BSAR_DEM_SOURCE_ASCII_GRID = 1, BSAR_DEM_SOURCE_GEOTIFF = 2 };

BSAR_API int bsar_dem_load(const char* path, int mode, bsar_dem_grid* out);
BSAR_API void bsar_dem_grid_free(bsar_dem_grid* g);

/* Geodesy of the GeoTIFF reader (WGS84; no counterpart in the reference,
 * whose GDAL loader documents the conversion, load_dem.h:9-17): UTM
 * easting/northing of (lat, lon) in degrees in `zone` (0: the zone of lon),
 * Krueger's series to fourth order in n; *zone_out is negative south of the
 * equator.  bsar_meridian_arc: metres from the equator to lat. */
BSAR_API int bsar_geo_to_utm(double lat, double lon, int zone, double* easting, double* northing, int* zone_out);
BSAR_API double bsar_meridian_arc(double lat);

/* ---- accounting --------------------------------------------------------- */

/* Algorithmic FP64 flop per voxel-pulse of the plain form: two 709-candidate
 * refraction solves (46 flop per candidate + 64 set-up), R1 means (20 N),
 * delay/phase, lerp and accumulate (48): 65,406 + 20 N. */
BSAR_API double bsar_flops_per_voxel_pulse(int n_pulses);

/* ---- serving loop --------------------------------------------------------
 *
 * No counterpart in the reference, whose driver images one grid per process
 * run (proc_test_script.cpp:main).  A server owns one device-side scheduler:
 * a host thread admits submitted jobs in arrival order while slots and the
 * HBM cap allow, and runs scheduling rounds in which the admitted jobs'
 * (pulse group, 256-voxel tile, pulse window) work items are pulled by
 * persistent kernels from one list, so images of different jobs run side by
 * side.  A job's image is bit-identical to bsar_backproject of the same grid
 * over a dataset holding the job's first K pulses, whatever else runs. */

typedef struct bsar_server bsar_server;

typedef struct {
    int max_jobs;            /* jobs admitted at once; 0 = 256 */
    int reserved0;
    int64_t hbm_cap_bytes;   /* cap on the admitted jobs' device working memory; 0 = 200e9 */
    double round_seconds;    /* target device time of one scheduling round; 0 = 0.04 */
    double reserved[4];
} bsar_server_config;

typedef struct {
    const bsar_dataset* ds;  /* resident dataset on the server's device, range data loaded */
    int pulses;              /* the job images the first K pulses of ds; 0 = all */
    int precision;           /* BSAR_FP64 or BSAR_MIXED */
    const double* gx; int nx;
    const double* gy; int ny;
    const double* gz; int nz;
    double* out;             /* host image [nz][ny][nx] complex, written before the job is done */
} bsar_job_desc;

typedef enum { BSAR_JOB_QUEUED = 0, BSAR_JOB_RUNNING = 1, BSAR_JOB_DONE = 2, BSAR_JOB_FAILED = 3 } bsar_job_state;

typedef struct {
    int64_t id;
    int state;               /* bsar_job_state */
    int status;              /* BSAR_OK, or why the job failed */
    double t_submit, t_admit, t_done;   /* seconds on the server clock (bsar_server_clock) */
    double voxel_pulses;
    double r1_exact_frac;    /* mixed: share of R1 means summed exactly */
    int64_t footprint_bytes; /* device working memory held while admitted */
} bsar_job_info;

typedef struct {
    int64_t submitted, completed, failed, rejected;
    int64_t rounds, items;
    double device_seconds;   /* summed device time of the rounds */
    double voxel_pulses;     /* work of the completed jobs */
    int64_t hbm_in_use, hbm_peak;
    int active, queued;
} bsar_server_stats;

/* Starts the scheduler thread on `device`.  cfg may be null (defaults). */
BSAR_API int bsar_server_create(int device, const bsar_server_config* cfg, bsar_server** out);
/* Stops after the admitted and queued jobs finish; frees the server. */
BSAR_API void bsar_server_destroy(bsar_server* s);
BSAR_API const char* bsar_server_last_error(const bsar_server* s);
/* Seconds since the server started (steady clock). */
BSAR_API double bsar_server_clock(const bsar_server* s);
/* Queues a job (grid axes are copied; ds and out must stay valid until the
 * job is collected).  BSAR_ERR_CAPACITY when its working memory alone exceeds
 * the HBM cap. */
BSAR_API int bsar_server_submit(bsar_server* s, const bsar_job_desc* job, int64_t* id);
/* Non-blocking state of a job. */
BSAR_API int bsar_server_poll(bsar_server* s, int64_t id, bsar_job_info* info);
/* Blocks until the job is done or failed (timeout_s < 0: no limit), then
 * collects it (its id becomes unknown).  Returns the job's status, or
 * BSAR_ERR_STATE on timeout. */
BSAR_API int bsar_server_wait(bsar_server* s, int64_t id, double timeout_s, bsar_job_info* info);
/* Blocks until any uncollected job is done or failed, collects it and
 * reports it in *info (info->id names it).  Returns that job's status,
 * BSAR_ERR_STATE on timeout, BSAR_ERR_INVALID when no job is uncollected. */
BSAR_API int bsar_server_wait_any(bsar_server* s, double timeout_s, bsar_job_info* info);
/* Blocks until no job is queued or running. */
BSAR_API int bsar_server_drain(bsar_server* s);
BSAR_API int bsar_server_get_stats(bsar_server* s, bsar_server_stats* st);

#ifdef __cplusplus
}
#endif
#endif /* BSAR_H */
