/*
 * skelcl_rt.h -- C-ABI runtime of the MI355X skeleton library.
 *
 * Replaces the reference's OpenCL runtime layer for one GPU:
 *   Device / DeviceList    include/SkelCL/detail/Device.h:123-245,
 *                          include/SkelCL/detail/DeviceList.h:62-122,
 *                          src/Device.cpp:95-341, src/DeviceList.cpp:82-173
 *   DeviceBuffer           include/SkelCL/detail/DeviceBuffer.h:57-98,
 *                          src/DeviceBuffer.cpp:84-174
 *   Program / Source       include/SkelCL/detail/Program.h:61-114,
 *                          src/Program.cpp:83-281, src/Source.cpp:56-84
 *   KernelUtil             include/SkelCL/detail/KernelUtil.h:134-355
 *   util::hash (SHA1)      src/Util.cpp:159-180
 *
 * Conventions: every function returns an int status (SKL_OK = 0); on error
 * skl_last_error() returns a thread-local message.  Handles are opaque.
 * Device memory is addressed by plain device pointers.  Offsets and sizes are
 * in BYTES unless a parameter name says otherwise.  Nothing here takes or
 * returns a C++ or torch type.
 */
#ifndef SKELCL_RT_H
#define SKELCL_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum skl_status {
  SKL_OK = 0,
  SKL_ERR_INVALID = 1,     /* bad argument / precondition (reference ASSERT) */
  SKL_ERR_NO_DEVICE = 2,   /* no GPU or bad device id (DeviceList.cpp:131-155) */
  SKL_ERR_HIP = 3,         /* HIP runtime error */
  SKL_ERR_COMPILE = 4,     /* hiprtc build failure (Program.cpp:202-210) */
  SKL_ERR_PARSE = 5,       /* user source could not be parsed */
  SKL_ERR_UNSUPPORTED = 6, /* feature not available on this path */
  SKL_ERR_NOT_INIT = 7     /* skl_init not called */
};

enum skl_device_type { SKL_DEVICE_ANY = 0, SKL_DEVICE_GPU = 1 };

typedef struct skl_stream_s* skl_stream;   /* NULL = the library stream */
typedef struct skl_program_s* skl_program;
typedef struct skl_kernel_s* skl_kernel;
typedef struct skl_event_s* skl_event;

typedef struct {
  char name[256];
  char arch[64];            /* e.g. "gfx950:sramecc+:xnack-" */
  int compute_units;
  int wavefront_size;
  int max_work_group_size;
  size_t lds_per_block;     /* bytes */
  size_t global_mem;        /* bytes */
  int clock_khz;
  int memory_clock_khz;
  int memory_bus_width;
  int supports_double;
  int l2_cache_bytes;
} skl_device_info;

/* ---- errors ------------------------------------------------------------ */
const char* skl_last_error(void);
void skl_free_string(char* s);

/* ---- DeviceList::init / terminate (src/SkelCL.cpp:51-85,
 *      src/DeviceList.cpp:82-157).  ndevices <= 0 means "all"; on this
 *      build every request collapses to GPU 0. ----------------------------- */
int skl_init(int ndevices, int device_type);
int skl_terminate(void);
int skl_is_initialized(void);
int skl_device_count(int* count);
int skl_device_get_info(int device, skl_device_info* info);
/* DeviceList::barrier (DeviceList.h:62-122): wait for all queued work */
int skl_device_synchronize(void);

/* ---- queues (Device's in-order queue, Device.cpp:95-122) ------------------ */
int skl_stream_create(skl_stream* stream);
int skl_stream_destroy(skl_stream stream);
int skl_stream_synchronize(skl_stream stream);
/* keepAlive / completion callbacks (Device.cpp:78-87, KernelUtil.h:288-355) */
int skl_stream_add_callback(skl_stream stream, void (*fn)(void* user),
                            void* user);

/* ---- DeviceBuffer (DeviceBuffer.cpp:84-174) ------------------------------ */
int skl_malloc(void** dptr, size_t bytes);
int skl_free(void* dptr);
int skl_host_malloc(void** hptr, size_t bytes); /* pinned host memory */
int skl_host_free(void* hptr);
/* Device::enqueueWrite (Device.cpp:169-199) */
int skl_memcpy_h2d(void* dst, size_t dst_offset, const void* src,
                   size_t bytes, skl_stream stream, int blocking);
/* Device::enqueueRead (Device.cpp:201-299) */
int skl_memcpy_d2h(void* dst, const void* src, size_t src_offset,
                   size_t bytes, skl_stream stream, int blocking);
/* Device::enqueueCopy (Device.cpp:301-330) */
int skl_memcpy_d2d(void* dst, size_t dst_offset, const void* src,
                   size_t src_offset, size_t bytes, skl_stream stream);
/* fill `count` elements of `elem_size` bytes (1,2,4,8,16) with `pattern`;
 * a kernel, so it may be captured into a graph */
int skl_fill(void* dst, const void* pattern, size_t elem_size, size_t count,
             skl_stream stream);

/* ---- events (Event.cpp:89-98; profiling as executor/executor.cpp:7-24) ---- */
int skl_event_create(skl_event* ev);
int skl_event_destroy(skl_event ev);
int skl_event_record(skl_event ev, skl_stream stream);
int skl_event_synchronize(skl_event ev);
int skl_event_elapsed_ms(skl_event start, skl_event stop, float* ms);

/* ---- Program (Program.cpp:83-281) ----------------------------------------
 * skl_program_compile: hiprtc compile of `source` for the target arch
 *   (gfx950 unless SKELCL_ARCH overrides); needs no GPU.  *code is malloc'd
 *   (free with skl_free_string), *log likewise (may be NULL).
 * skl_program_build: compile and load on GPU 0; a program built before in
 *   this process is shared (in-memory cache).
 * Both use the on-disk code-object cache only as the reference's binary
 *   cache does (Program.cpp:139,256): a code object is read when
 *   SKELCL_LOAD_BINARY=YES and written when SKELCL_SAVE_BINARY=YES, as
 *   $SKELCL_CACHE_DIR (else ~/.cache/skelcl_amd) / SHA1(source | options |
 *   arch | hiprtc version).co; a cached object the runtime rejects is
 *   compiled afresh.  SKELCL_PRINT_BUILD_LOG=YES prints the hiprtc log
 *   (Program.cpp:198). */
int skl_program_compile(const char* source, const char* options, char** code,
                        size_t* code_size, char** log);
int skl_program_build(const char* source, const char* options,
                      skl_program* program, char** log);
int skl_program_release(skl_program program);
int skl_kernel_get(skl_program program, const char* name, skl_kernel* kernel);
/* Device::enqueue (Device.cpp:124-167) with kernelUtil::setKernelArgs
 * (KernelUtil.h:134-245): args[i] points at the i-th kernel argument value. */
int skl_kernel_launch(skl_kernel kernel, const unsigned int grid[3],
                      const unsigned int block[3], size_t dynamic_lds,
                      void** args, skl_stream stream);

/* ---- util::hash (src/Util.cpp:159-180): 40 hex chars + NUL ----------------- */
int skl_sha1_hex(const void* data, size_t len, char out[41]);

/* ---- graphs: capture the work enqueued on `stream` between begin and end -- */
typedef struct skl_graph_s* skl_graph;
int skl_graph_begin_capture(skl_stream stream);
int skl_graph_end_capture(skl_stream stream, skl_graph* graph);
int skl_graph_launch(skl_graph graph, skl_stream stream);
int skl_graph_destroy(skl_graph graph);

/* ---- synthetic inputs (counter-based SplitMix64, bit-identical to a host
 * implementation of the same formulas) --------------------------------------
 *   z = seed + 0x9E3779B97F4A7C15*(i+1); z = (z^z>>30)*0xBF58476D1CE4E5B9;
 *   z = (z^z>>27)*0x94D049BB133111EB; z ^= z>>31
 *   f32 symmetric:  2*((z>>40)*2^-24) - 1        (uniform [-1,1))
 *   f32 unit:       scale*((z>>40)*2^-24)
 *   i32 below m:    ((z>>32)*m)>>32              (uniform [0,m)) */
int skl_generate_f32_sym(float* out, size_t n, uint64_t seed,
                         skl_stream stream);
int skl_generate_f32_unit(float* out, size_t n, uint64_t seed, float scale,
                          skl_stream stream);
int skl_generate_i32_below(int32_t* out, size_t n, uint64_t seed, uint32_t m,
                           skl_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* SKELCL_RT_H */
