/* sfa_io.h — reader/writer for the SFA field-archive container used by the
 * sfa/sim snapshots (format: reference sfa/format/SFA_SPEC.md, byte layout as
 * written by sfa/format/sfa.h). Own implementation; zstd is loaded at run time
 * with dlopen so RAW archives work on hosts without libzstd.
 *
 * Layout kept identical to the reference writer, quirks included:
 *   - SFAH chunk size field = 100 (sfa.h:649-650), total_frames at offset 68,
 *     streaming flag cleared on close (sfa.h:1736-1757);
 *   - KVMD size field = payload length, excluding the 12-byte header (sfa.h:717-718);
 *   - FRMD size field includes its header (sfa.h:993-995);
 *   - JTOP/JMPF pre-allocated with 1024 zeroed slots, JMPF entries are 32 bytes
 *     {time, offset, comp_size, crc32c, frame_type} (sfa.h:998-1005);
 *   - COLZSTD payload = [n_cols u32][comp sizes u64 x n][BSS+zstd(3) per column].
 */
#ifndef SCP_SFA_IO_H
#define SCP_SFA_IO_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SFAIO_F16 = 0, SFAIO_F32 = 1, SFAIO_F64 = 2 };
enum { SFAIO_POSITION = 0, SFAIO_ANGLE = 1, SFAIO_VELOCITY = 2 };
enum { SFAIO_CODEC_RAW = 0, SFAIO_CODEC_ZSTD = 1, SFAIO_CODEC_BSS = 2, SFAIO_CODEC_COLZSTD = 7 };
/* JMPF frame types (sfa.h:73-76) */
enum { SFAIO_FRAME_VOXEL = 0, SFAIO_FRAME_VEC_I = 1, SFAIO_FRAME_VEC_P = 2 };
#define SFAIO_FLAG_STREAMING (1u << 5)
#define SFAIO_MAX_COLS 64
#define SFAIO_MAX_KVMD 16
#define SFAIO_MAX_PAIRS 128

typedef struct {
    char    name[12];
    uint8_t dtype, semantic, component, flags;
    double  scale;
} sfaio_col;

typedef struct {
    uint16_t set_id;
    uint32_t first_frame, frame_count;
    int      n_pairs;
    char     keys[SFAIO_MAX_PAIRS][64];
    char     vals[SFAIO_MAX_PAIRS][256];
} sfaio_kvmd;

typedef struct sfaio {
    FILE    *fp;
    int      writing;
    uint32_t version, flags;
    uint32_t Nx, Ny, Nz;
    double   Lx, Ly, Lz, dt;
    uint32_t n_columns, total_frames;
    uint64_t jtop_offset, cdef_offset;
    uint32_t jtop_max, jmpf_max;
    uint64_t n_total, frame_bytes;
    sfaio_col cols[SFAIO_MAX_COLS];
    /* writer state */
    uint64_t cur_jmpf_offset;
    uint32_t cur_jmpf_entries, cur_l1;
    int      n_kvmd_pending;
    uint8_t *kvmd_payload[SFAIO_MAX_KVMD];
    uint64_t kvmd_len[SFAIO_MAX_KVMD];
    /* reader state */
    int      n_kvmd;
    sfaio_kvmd *kvmd;
} sfaio;

extern const int sfaio_dtype_size[12];

uint32_t sfaio_crc32c(const void *data, size_t len);
int      sfaio_zstd_available(void);

sfaio *sfaio_create(const char *path, uint32_t Nx, uint32_t Ny, uint32_t Nz,
                    double Lx, double Ly, double Lz, double dt);
int    sfaio_set_codec(sfaio *s, int codec, int streaming);
int    sfaio_add_column(sfaio *s, const char *name, int dtype, int semantic, int component);
int    sfaio_add_kvmd(sfaio *s, uint16_t set_id, uint32_t first_frame, uint32_t frame_count,
                      const char *const *keys, const char *const *vals, int n_pairs);
int    sfaio_finalize_header(sfaio *s);
int    sfaio_write_frame(sfaio *s, double time, const void *const *cols);
/* Vector frame payload (the FRVD I/P layouts of sfa.h:1643-1720), zstd level 3. */
int    sfaio_write_vec(sfaio *s, double time, uint32_t frame_type, const void *payload, uint64_t len);
int    sfaio_close(sfaio *s);

sfaio *sfaio_open(const char *path);
/* Voxel frames only: a vector frame returns -4. */
int    sfaio_read_frame(sfaio *s, uint32_t frame, void *buf);
/* Frame type from the index (SFAIO_FRAME_*), -1 if out of range. */
int    sfaio_frame_type(sfaio *s, uint32_t frame);
/* Decompressed payload of a vector frame into buf (capacity cap); its length, or < 0. */
int64_t sfaio_read_vec(sfaio *s, uint32_t frame, void *buf, uint64_t cap);
double sfaio_frame_time(sfaio *s, uint32_t frame);
/* Convenience getters for FFI callers that do not map the struct. */
void   sfaio_info(const sfaio *s, uint32_t u[6], double d[4]);
int    sfaio_column_info(const sfaio *s, int c, char name[12], uint8_t meta[4]);
int    sfaio_kvmd_count(const sfaio *s);
int    sfaio_kvmd_pair(const sfaio *s, int set, int pair, const char **key, const char **val,
                       uint32_t *first_frame);
/* Streaming repair: index FRMD and FRVD chunks whose JMPF slots are zero (a writer
 * that died mid-run), patch total_frames, clear the streaming flag.
 * Returns the number of indexed frames or -1. */
int    sfaio_fixup_index(const char *path);

#ifdef __cplusplus
}
#endif
#endif
