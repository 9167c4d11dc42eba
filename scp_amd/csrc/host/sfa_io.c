/* sfa_io.c — SFA container reader/writer (see sfa_io.h for the layout contract). */
#define _FILE_OFFSET_BITS 64
#define _GNU_SOURCE
#include "sfa_io.h"

#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

const int sfaio_dtype_size[12] = {2, 4, 8, 16, 1, 2, 4, 8, 1, 2, 4, 8};

/* ---------------- CRC32C (Castagnoli, reflected 0x82F63B78), slicing by 8 ---------------- */

static uint32_t crc_tab[8][256];
static int crc_ready = 0;

static void crc_init(void) {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int b = 0; b < 8; b++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
        crc_tab[0][n] = c;
    }
    for (uint32_t n = 0; n < 256; n++)
        for (int t = 1; t < 8; t++)
            crc_tab[t][n] = (crc_tab[t - 1][n] >> 8) ^ crc_tab[0][crc_tab[t - 1][n] & 0xFF];
    crc_ready = 1;
}

uint32_t sfaio_crc32c(const void *data, size_t len) {
    if (!crc_ready) crc_init();
    const uint8_t *p = (const uint8_t *)data;
    uint32_t c = 0xFFFFFFFFu;
    while (len >= 8) {
        uint64_t w; memcpy(&w, p, 8);
        uint32_t lo = (uint32_t)w ^ c, hi = (uint32_t)(w >> 32);
        c = crc_tab[7][lo & 0xFF] ^ crc_tab[6][(lo >> 8) & 0xFF] ^ crc_tab[5][(lo >> 16) & 0xFF] ^ crc_tab[4][lo >> 24]
          ^ crc_tab[3][hi & 0xFF] ^ crc_tab[2][(hi >> 8) & 0xFF] ^ crc_tab[1][(hi >> 16) & 0xFF] ^ crc_tab[0][hi >> 24];
        p += 8; len -= 8;
    }
    while (len--) c = (c >> 8) ^ crc_tab[0][(c ^ *p++) & 0xFF];
    return ~c;
}

/* ---------------- zstd, resolved at run time ---------------- */

typedef size_t (*zc_fn)(void *, size_t, const void *, size_t, int);
typedef size_t (*zd_fn)(void *, size_t, const void *, size_t);
typedef size_t (*zb_fn)(size_t);
typedef unsigned (*ze_fn)(size_t);
static struct { int tried; zc_fn comp; zd_fn decomp; zb_fn bound; ze_fn iserr; } Z;

static int zstd_load(void) {
    if (Z.tried) return Z.comp != NULL;
    Z.tried = 1;
    const char *cands[3] = { getenv("SCP_ZSTD_LIB"), "libzstd.so.1", "/opt/conda/lib/libzstd.so.1" };
    for (int i = 0; i < 3; i++) {
        if (!cands[i] || !cands[i][0]) continue;
        void *h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        Z.comp = (zc_fn)dlsym(h, "ZSTD_compress");
        Z.decomp = (zd_fn)dlsym(h, "ZSTD_decompress");
        Z.bound = (zb_fn)dlsym(h, "ZSTD_compressBound");
        Z.iserr = (ze_fn)dlsym(h, "ZSTD_isError");
        if (Z.comp && Z.decomp && Z.bound && Z.iserr) return 1;
        Z.comp = NULL;
    }
    return 0;
}

int sfaio_zstd_available(void) { return zstd_load(); }

/* byte-stream split: value-major -> byte-plane-major and back */
static void bss_split(const uint8_t *src, uint8_t *dst, uint64_t n, int es) {
    for (int b = 0; b < es; b++) {
        uint8_t *d = dst + (uint64_t)b * n; const uint8_t *s = src + b;
        for (uint64_t i = 0; i < n; i++) d[i] = s[i * es];
    }
}
static void bss_join(const uint8_t *src, uint8_t *dst, uint64_t n, int es) {
    for (int b = 0; b < es; b++) {
        const uint8_t *s = src + (uint64_t)b * n; uint8_t *d = dst + b;
        for (uint64_t i = 0; i < n; i++) d[i * es] = s[i];
    }
}

/* ---------------- little-endian field helpers ---------------- */

static int wr(FILE *fp, const void *p, size_t n) { return fwrite(p, 1, n, fp) == n ? 0 : -1; }
static int rd(FILE *fp, void *p, size_t n) { return fread(p, 1, n, fp) == n ? 0 : -1; }
static int chunk_hdr(FILE *fp, const char *t, uint64_t size) { return wr(fp, t, 4) | wr(fp, &size, 8); }
static int seek(FILE *fp, uint64_t off) { return fseeko(fp, (off_t)off, SEEK_SET); }
static int wr_at(FILE *fp, uint64_t off, const void *p, size_t n) { return seek(fp, off) | wr(fp, p, n); }

static void zeros(FILE *fp, uint64_t n) {
    static const uint8_t z[4096];
    while (n) { size_t k = n > sizeof z ? sizeof z : (size_t)n; fwrite(z, 1, k, fp); n -= k; }
}

static void update_sizes(sfaio *s) {
    s->n_total = (uint64_t)s->Nx * s->Ny * s->Nz;
    s->frame_bytes = 0;
    for (uint32_t c = 0; c < s->n_columns; c++)
        s->frame_bytes += s->n_total * (uint64_t)sfaio_dtype_size[s->cols[c].dtype];
}

/* ---------------- writer ---------------- */

sfaio *sfaio_create(const char *path, uint32_t Nx, uint32_t Ny, uint32_t Nz,
                    double Lx, double Ly, double Lz, double dt) {
    FILE *fp = fopen(path, "w+b");
    if (!fp) return NULL;
    sfaio *s = calloc(1, sizeof *s);
    s->fp = fp; s->writing = 1; s->version = 3;
    s->flags = SFAIO_CODEC_BSS | SFAIO_FLAG_STREAMING;     /* sfa.h:574 default */
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz; s->Lx = Lx; s->Ly = Ly; s->Lz = Lz; s->dt = dt;
    s->jtop_max = 1024; s->jmpf_max = 1024;
    return s;
}

int sfaio_set_codec(sfaio *s, int codec, int streaming) {
    if (codec != SFAIO_CODEC_RAW && codec != SFAIO_CODEC_ZSTD && codec != SFAIO_CODEC_BSS
        && codec != SFAIO_CODEC_COLZSTD) return -1;
    if (codec != SFAIO_CODEC_RAW && !zstd_load()) return -2;
    s->flags = (uint32_t)codec | (streaming ? SFAIO_FLAG_STREAMING : 0);
    return 0;
}

int sfaio_add_column(sfaio *s, const char *name, int dtype, int semantic, int component) {
    if (s->n_columns >= SFAIO_MAX_COLS || dtype < 0 || dtype > 11) return -1;
    sfaio_col *c = &s->cols[s->n_columns++];
    memset(c, 0, sizeof *c);
    strncpy(c->name, name, 11);
    c->dtype = (uint8_t)dtype; c->semantic = (uint8_t)semantic; c->component = (uint8_t)component;
    c->scale = 1.0;
    return 0;
}

int sfaio_add_kvmd(sfaio *s, uint16_t set_id, uint32_t first_frame, uint32_t frame_count,
                   const char *const *keys, const char *const *vals, int n_pairs) {
    if (s->n_kvmd_pending >= SFAIO_MAX_KVMD || n_pairs < 0 || n_pairs > 65535) return -1;
    uint64_t len = 16;
    for (int i = 0; i < n_pairs; i++) len += strlen(keys[i]) + 1 + strlen(vals[i]) + 1;
    uint8_t *b = calloc(1, len);
    uint16_t np = (uint16_t)n_pairs;
    b[0] = 1;                                   /* sub-header version */
    memcpy(b + 1, &set_id, 2); memcpy(b + 3, &np, 2);
    memcpy(b + 5, &first_frame, 4); memcpy(b + 9, &frame_count, 4);
    uint64_t o = 16;
    for (int i = 0; i < n_pairs; i++) {
        size_t kl = strlen(keys[i]) + 1, vl = strlen(vals[i]) + 1;
        memcpy(b + o, keys[i], kl); o += kl; memcpy(b + o, vals[i], vl); o += vl;
    }
    s->kvmd_payload[s->n_kvmd_pending] = b; s->kvmd_len[s->n_kvmd_pending] = len;
    s->n_kvmd_pending++;
    return 0;
}

static void write_jmpf(FILE *fp, uint32_t jmpf_max) {
    uint32_t zero = 0;
    chunk_hdr(fp, "JMPF", 12 + 8 + (uint64_t)jmpf_max * 32);
    wr(fp, &jmpf_max, 4); wr(fp, &zero, 4);
    zeros(fp, (uint64_t)jmpf_max * 32);
}

int sfaio_finalize_header(sfaio *s) {
    FILE *fp = s->fp;
    update_sizes(s);
    uint64_t zero64 = 0; uint32_t zero32 = 0;
    if (seek(fp, 0)) return -1;
    chunk_hdr(fp, "SFAH", 100);
    wr(fp, &s->version, 4); wr(fp, &s->flags, 4);
    wr(fp, &s->Nx, 4); wr(fp, &s->Ny, 4); wr(fp, &s->Nz, 4);
    wr(fp, &s->Lx, 8); wr(fp, &s->Ly, 8); wr(fp, &s->Lz, 8); wr(fp, &s->dt, 8);
    wr(fp, &s->n_columns, 4); wr(fp, &zero32, 4);          /* total_frames, patched at close */
    wr(fp, &zero64, 8); wr(fp, &zero64, 8);                /* jtop / cdef offsets, patched below */
    wr(fp, &s->jtop_max, 4); wr(fp, &s->jmpf_max, 4); wr(fp, &zero32, 4);
    s->cdef_offset = (uint64_t)ftello(fp);
    chunk_hdr(fp, "CDEF", 12 + (uint64_t)s->n_columns * 24);
    for (uint32_t c = 0; c < s->n_columns; c++) {
        sfaio_col *k = &s->cols[c];
        wr(fp, k->name, 12); wr(fp, &k->dtype, 1); wr(fp, &k->semantic, 1);
        wr(fp, &k->component, 1); wr(fp, &k->flags, 1); wr(fp, &k->scale, 8);
    }
    for (int i = 0; i < s->n_kvmd_pending; i++) {
        chunk_hdr(fp, "KVMD", s->kvmd_len[i]);             /* size excludes the header (sfa.h:717) */
        wr(fp, s->kvmd_payload[i], s->kvmd_len[i]);
        free(s->kvmd_payload[i]); s->kvmd_payload[i] = NULL;
    }
    s->n_kvmd_pending = 0;
    s->jtop_offset = (uint64_t)ftello(fp);
    chunk_hdr(fp, "JTOP", 12 + 16 + (uint64_t)s->jtop_max * 16);
    uint32_t one = 1;
    wr(fp, &s->jtop_max, 4); wr(fp, &one, 4); wr(fp, &zero64, 8);
    s->cur_jmpf_offset = s->jtop_offset + 12 + 16 + (uint64_t)s->jtop_max * 16;
    wr(fp, &s->cur_jmpf_offset, 8); wr(fp, &zero32, 4); wr(fp, &zero32, 4);  /* L1 entry 0 */
    zeros(fp, (uint64_t)(s->jtop_max - 1) * 16);
    write_jmpf(fp, s->jmpf_max);
    s->cur_jmpf_entries = 0; s->cur_l1 = 0;
    wr_at(fp, 12 + 4 + 4 + 12 + 24 + 8 + 4 + 4, &s->jtop_offset, 8);   /* offset 60 */
    wr_at(fp, 12 + 4 + 4 + 12 + 24 + 8 + 4 + 4 + 8, &s->cdef_offset, 8);
    if (fflush(fp)) return -1;
    return fseeko(fp, 0, SEEK_END);
}

/* Encode one frame into a payload according to the codec. */
static uint8_t *encode_frame(sfaio *s, const uint8_t *raw, uint64_t *out_len) {
    int codec = s->flags & 0xF;
    if (codec == SFAIO_CODEC_RAW) {
        uint8_t *p = malloc(s->frame_bytes); memcpy(p, raw, s->frame_bytes);
        *out_len = s->frame_bytes; return p;
    }
    if (!zstd_load()) return NULL;
    if (codec == SFAIO_CODEC_COLZSTD) {
        uint32_t nc = s->n_columns;
        uint8_t **cd = calloc(nc, sizeof *cd); uint64_t *cs = calloc(nc, sizeof *cs);
        uint64_t off = 0, total = 4 + 8ull * nc; int err = 0;
        for (uint32_t c = 0; c < nc && !err; c++) {
            int es = sfaio_dtype_size[s->cols[c].dtype];
            uint64_t cb = s->n_total * es;
            uint8_t *tmp = malloc(cb);
            bss_split(raw + off, tmp, s->n_total, es);
            size_t bound = Z.bound(cb);
            cd[c] = malloc(bound);
            size_t r = Z.comp(cd[c], bound, tmp, cb, 3);
            free(tmp);
            if (Z.iserr(r)) err = 1; else { cs[c] = r; total += r; }
            off += cb;
        }
        uint8_t *p = NULL;
        if (!err) {
            p = malloc(total);
            memcpy(p, &nc, 4); memcpy(p + 4, cs, 8ull * nc);
            uint64_t o = 4 + 8ull * nc;
            for (uint32_t c = 0; c < nc; c++) { memcpy(p + o, cd[c], cs[c]); o += cs[c]; }
            *out_len = total;
        }
        for (uint32_t c = 0; c < nc; c++) free(cd[c]);
        free(cd); free(cs);
        return p;
    }
    const uint8_t *src = raw; uint8_t *tmp = NULL;
    if (codec == SFAIO_CODEC_BSS) {
        tmp = malloc(s->frame_bytes);
        uint64_t off = 0;
        for (uint32_t c = 0; c < s->n_columns; c++) {
            int es = sfaio_dtype_size[s->cols[c].dtype];
            bss_split(raw + off, tmp + off, s->n_total, es);
            off += s->n_total * es;
        }
        src = tmp;
    }
    size_t bound = Z.bound(s->frame_bytes);
    uint8_t *p = malloc(bound);
    size_t r = Z.comp(p, bound, src, s->frame_bytes, 3);
    free(tmp);
    if (Z.iserr(r)) { free(p); return NULL; }
    *out_len = r;
    return p;
}

/* Before appending a frame: open a new L2 table (JMPF) when the current one is full. */
static void open_slot(sfaio *s) {
    FILE *fp = s->fp;
    uint32_t zero32 = 0;
    if (s->cur_jmpf_entries < s->jmpf_max) return;
    fseeko(fp, 0, SEEK_END);
    uint64_t nj = (uint64_t)ftello(fp);
    write_jmpf(fp, s->jmpf_max);
    s->cur_l1++;
    uint64_t ent = s->jtop_offset + 12 + 16 + (uint64_t)s->cur_l1 * 16;
    wr_at(fp, ent, &nj, 8); wr(fp, &s->total_frames, 4); wr(fp, &zero32, 4);
    uint32_t ne = s->cur_l1 + 1;
    wr_at(fp, s->jtop_offset + 12 + 4, &ne, 4);
    s->cur_jmpf_offset = nj; s->cur_jmpf_entries = 0;
}

/* JMPF entry of the frame just appended: time, chunk offset, payload size, CRC, type. */
static int index_frame(sfaio *s, double time, uint64_t off, uint64_t size, uint32_t crc, uint32_t type) {
    FILE *fp = s->fp;
    uint64_t slot = s->cur_jmpf_offset + 12 + 8 + (uint64_t)s->cur_jmpf_entries * 32;
    wr_at(fp, slot, &time, 8); wr(fp, &off, 8); wr(fp, &size, 8); wr(fp, &crc, 4); wr(fp, &type, 4);
    s->cur_jmpf_entries++;
    wr_at(fp, s->cur_jmpf_offset + 12 + 4, &s->cur_jmpf_entries, 4);
    wr_at(fp, s->jtop_offset + 12 + 16 + (uint64_t)s->cur_l1 * 16 + 12, &s->cur_jmpf_entries, 4);
    s->total_frames++;
    fflush(fp);
    return fseeko(fp, 0, SEEK_END);
}

int sfaio_write_frame(sfaio *s, double time, const void *const *cols) {
    if (!s->writing) return -1;
    FILE *fp = s->fp;
    uint8_t *raw = malloc(s->frame_bytes);
    if (!raw) return -1;
    uint64_t off = 0;
    for (uint32_t c = 0; c < s->n_columns; c++) {
        uint64_t cb = s->n_total * sfaio_dtype_size[s->cols[c].dtype];
        memcpy(raw + off, cols[c], cb); off += cb;
    }
    uint32_t crc = sfaio_crc32c(raw, s->frame_bytes);      /* on the untransformed frame */
    uint64_t plen = 0;
    uint8_t *payload = encode_frame(s, raw, &plen);
    free(raw);
    if (!payload) return -1;
    open_slot(s);
    fseeko(fp, 0, SEEK_END);
    uint64_t frmd = (uint64_t)ftello(fp);
    int e = chunk_hdr(fp, "FRMD", 12 + plen) | wr(fp, payload, plen);
    free(payload);
    if (e) return -1;
    return index_frame(s, time, frmd, plen, crc, SFAIO_FRAME_VOXEL);
}

/* Vector frame (FRVD chunk, sfa.h:1563-1640): type(4), size(8) = 8 + the
 * compressed length, the frame time(8), the zstd-compressed payload; indexed
 * like a voxel frame with its frame type and the CRC32C of the payload. */
int sfaio_write_vec(sfaio *s, double time, uint32_t frame_type, const void *payload, uint64_t len) {
    if (!s->writing || !zstd_load()) return -1;
    FILE *fp = s->fp;
    size_t bound = Z.bound(len);
    uint8_t *comp = malloc(bound);
    if (!comp) return -1;
    size_t clen = Z.comp(comp, bound, payload, len, 3);
    if (Z.iserr(clen)) { free(comp); return -1; }
    uint32_t crc = sfaio_crc32c(payload, len);
    open_slot(s);
    fseeko(fp, 0, SEEK_END);
    uint64_t off = (uint64_t)ftello(fp);
    uint64_t c64 = clen;
    int e = chunk_hdr(fp, "FRVD", 8 + c64) | wr(fp, &time, 8) | wr(fp, comp, clen);
    free(comp);
    if (e) return -1;
    return index_frame(s, time, off, c64, crc, frame_type);
}

int sfaio_close(sfaio *s) {
    if (!s) return -1;
    int rc = 0;
    if (s->writing && s->fp) {
        rc |= wr_at(s->fp, 68, &s->total_frames, 4);
        uint32_t fl = s->flags & ~SFAIO_FLAG_STREAMING;
        rc |= wr_at(s->fp, 16, &fl, 4);
    }
    if (s->fp) rc |= fclose(s->fp);
    for (int i = 0; i < s->n_kvmd_pending; i++) free(s->kvmd_payload[i]);
    free(s->kvmd);
    free(s);
    return rc ? -1 : 0;
}

/* ---------------- reader ---------------- */

static void read_kvmd_sets(sfaio *s) {
    uint64_t pos = s->cdef_offset + 12 + (uint64_t)s->n_columns * 24;
    s->kvmd = calloc(SFAIO_MAX_KVMD, sizeof(sfaio_kvmd));
    s->n_kvmd = 0;
    while (pos + 12 <= s->jtop_offset && s->n_kvmd < SFAIO_MAX_KVMD) {
        char t[4]; uint64_t size;
        if (seek(s->fp, pos) || rd(s->fp, t, 4) || rd(s->fp, &size, 8)) break;
        if (memcmp(t, "KVMD", 4) == 0 && size >= 16 && size < 100000000) {
            uint8_t *b = malloc(size + 1);
            if (rd(s->fp, b, size)) { free(b); break; }
            b[size] = 0;
            sfaio_kvmd *k = &s->kvmd[s->n_kvmd++];
            uint16_t np;
            memcpy(&k->set_id, b + 1, 2); memcpy(&np, b + 3, 2);
            memcpy(&k->first_frame, b + 5, 4); memcpy(&k->frame_count, b + 9, 4);
            uint64_t o = 16;
            for (int i = 0; i < np && o < size && k->n_pairs < SFAIO_MAX_PAIRS; i++) {
                const char *key = (const char *)b + o; o += strlen(key) + 1;
                if (o >= size) break;
                const char *val = (const char *)b + o; o += strlen(val) + 1;
                strncpy(k->keys[k->n_pairs], key, 63);
                strncpy(k->vals[k->n_pairs], val, 255);
                k->n_pairs++;
            }
            free(b);
        }
        pos += 12 + size;   /* KVMD and GDEF sizes exclude the header */
    }
}

sfaio *sfaio_open(const char *path) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return NULL;
    sfaio *s = calloc(1, sizeof *s);
    s->fp = fp;
    char t[4]; uint64_t size; uint32_t reserved;
    int e = rd(fp, t, 4) | rd(fp, &size, 8);
    if (e || memcmp(t, "SFAH", 4)) { fclose(fp); free(s); return NULL; }
    e |= rd(fp, &s->version, 4) | rd(fp, &s->flags, 4);
    e |= rd(fp, &s->Nx, 4) | rd(fp, &s->Ny, 4) | rd(fp, &s->Nz, 4);
    e |= rd(fp, &s->Lx, 8) | rd(fp, &s->Ly, 8) | rd(fp, &s->Lz, 8) | rd(fp, &s->dt, 8);
    e |= rd(fp, &s->n_columns, 4) | rd(fp, &s->total_frames, 4);
    e |= rd(fp, &s->jtop_offset, 8) | rd(fp, &s->cdef_offset, 8);
    e |= rd(fp, &s->jtop_max, 4) | rd(fp, &s->jmpf_max, 4) | rd(fp, &reserved, 4);
    if (e || s->n_columns > SFAIO_MAX_COLS || s->jmpf_max == 0 || seek(fp, s->cdef_offset + 12)) {
        fclose(fp); free(s); return NULL;
    }
    for (uint32_t c = 0; c < s->n_columns; c++) {
        sfaio_col *k = &s->cols[c];
        e |= rd(fp, k->name, 12) | rd(fp, &k->dtype, 1) | rd(fp, &k->semantic, 1);
        e |= rd(fp, &k->component, 1) | rd(fp, &k->flags, 1) | rd(fp, &k->scale, 8);
        if (k->dtype > 11) e = 1;
    }
    if (e) { fclose(fp); free(s); return NULL; }
    update_sizes(s);
    read_kvmd_sets(s);
    return s;
}

typedef struct { double time; uint64_t offset, size; uint32_t crc, type; } l2ent;

static int find_frame(sfaio *s, uint32_t frame, l2ent *o) {
    uint64_t jt = s->jtop_offset;
    while (jt) {
        uint32_t mx, cur; uint64_t next;
        if (seek(s->fp, jt + 12) || rd(s->fp, &mx, 4) || rd(s->fp, &cur, 4) || rd(s->fp, &next, 8)) return -1;
        uint32_t l1 = frame / s->jmpf_max;
        if (l1 < cur) {
            uint64_t jm; uint32_t first, cnt;
            if (seek(s->fp, jt + 12 + 16 + (uint64_t)l1 * 16) || rd(s->fp, &jm, 8) || rd(s->fp, &first, 4)
                || rd(s->fp, &cnt, 4)) return -1;
            uint32_t l2 = frame % s->jmpf_max;
            if (l2 >= cnt) return -1;
            if (seek(s->fp, jm + 12 + 8 + (uint64_t)l2 * 32) || rd(s->fp, &o->time, 8) || rd(s->fp, &o->offset, 8)
                || rd(s->fp, &o->size, 8) || rd(s->fp, &o->crc, 4) || rd(s->fp, &o->type, 4)) return -1;
            return 0;
        }
        frame -= cur * s->jmpf_max;
        jt = next;
    }
    return -1;
}

static int decode_payload(const sfaio *s, int codec, const uint8_t *comp, uint64_t clen, uint8_t *buf) {
    if (codec == SFAIO_CODEC_RAW) {
        if (clen != s->frame_bytes) return -1;
        memcpy(buf, comp, clen); return 0;
    }
    if (!zstd_load()) return -3;
    if (codec == SFAIO_CODEC_COLZSTD) {
        uint32_t nc;
        if (clen < 4) return -1;
        memcpy(&nc, comp, 4);
        if (nc != s->n_columns || clen < 4 + 8ull * nc) return -1;
        const uint8_t *p = comp + 4 + 8ull * nc;
        uint64_t out = 0;
        for (uint32_t c = 0; c < nc; c++) {
            uint64_t cs; memcpy(&cs, comp + 4 + 8ull * c, 8);
            int es = sfaio_dtype_size[s->cols[c].dtype];
            uint64_t cb = s->n_total * es;
            if ((uint64_t)(p - comp) + cs > clen) return -1;
            uint8_t *tmp = malloc(cb);
            size_t r = Z.decomp(tmp, cb, p, cs);
            if (Z.iserr(r) || r != cb) { free(tmp); return -1; }
            bss_join(tmp, buf + out, s->n_total, es);
            free(tmp);
            p += cs; out += cb;
        }
        return 0;
    }
    uint8_t *dst = buf, *tmp = NULL;
    if (codec == SFAIO_CODEC_BSS) { tmp = malloc(s->frame_bytes); dst = tmp; }
    size_t r = Z.decomp(dst, s->frame_bytes, comp, clen);
    if (Z.iserr(r) || r != s->frame_bytes) { free(tmp); return -1; }
    if (tmp) {
        uint64_t off = 0;
        for (uint32_t c = 0; c < s->n_columns; c++) {
            int es = sfaio_dtype_size[s->cols[c].dtype];
            bss_join(tmp + off, buf + off, s->n_total, es);
            off += s->n_total * es;
        }
        free(tmp);
    }
    return 0;
}

int sfaio_read_frame(sfaio *s, uint32_t frame, void *buf) {
    l2ent e;
    if (find_frame(s, frame, &e)) return -1;
    if (e.type != SFAIO_FRAME_VOXEL) return -4;              /* a vector frame: sfaio_read_vec */
    if (e.size == 0 || e.size > 50000000000ULL) return -1;
    uint8_t *comp = malloc(e.size);
    if (!comp) return -1;
    if (seek(s->fp, e.offset + 12) || rd(s->fp, comp, e.size)) { free(comp); return -1; }
    int rc = decode_payload(s, s->flags & 0xF, comp, e.size, (uint8_t *)buf);
    free(comp);
    if (rc) return rc;
    /* Unlike the reference reader, COLZSTD frames are checksummed too. */
    if (sfaio_crc32c(buf, s->frame_bytes) != e.crc) return -2;
    return 0;
}

int sfaio_frame_type(sfaio *s, uint32_t frame) {
    l2ent e;
    return find_frame(s, frame, &e) ? -1 : (int)e.type;
}

int64_t sfaio_read_vec(sfaio *s, uint32_t frame, void *buf, uint64_t cap) {
    l2ent e;
    if (find_frame(s, frame, &e)) return -1;
    if (e.type != SFAIO_FRAME_VEC_I && e.type != SFAIO_FRAME_VEC_P) return -4;
    if (!zstd_load()) return -3;
    if (e.size == 0 || e.size > 50000000000ULL) return -1;
    uint8_t *comp = malloc(e.size);
    if (!comp) return -1;
    if (seek(s->fp, e.offset + 12 + 8) || rd(s->fp, comp, e.size)) { free(comp); return -1; }
    size_t r = Z.decomp(buf, cap, comp, e.size);
    free(comp);
    if (Z.iserr(r)) return -1;
    if (e.crc && sfaio_crc32c(buf, r) != e.crc) return -2;   /* an index rebuilt by fixup holds 0 */
    return (int64_t)r;
}

double sfaio_frame_time(sfaio *s, uint32_t frame) {
    l2ent e;
    return find_frame(s, frame, &e) ? -1.0 : e.time;
}

void sfaio_info(const sfaio *s, uint32_t u[6], double d[4]) {
    u[0] = s->Nx; u[1] = s->Ny; u[2] = s->Nz; u[3] = s->n_columns; u[4] = s->total_frames; u[5] = s->flags;
    d[0] = s->Lx; d[1] = s->Ly; d[2] = s->Lz; d[3] = s->dt;
}

int sfaio_column_info(const sfaio *s, int c, char name[12], uint8_t meta[4]) {
    if (c < 0 || (uint32_t)c >= s->n_columns) return -1;
    memcpy(name, s->cols[c].name, 12);
    meta[0] = s->cols[c].dtype; meta[1] = s->cols[c].semantic; meta[2] = s->cols[c].component; meta[3] = s->cols[c].flags;
    return 0;
}

int sfaio_kvmd_count(const sfaio *s) { return s->n_kvmd; }

int sfaio_kvmd_pair(const sfaio *s, int set, int pair, const char **key, const char **val, uint32_t *first_frame) {
    if (set < 0 || set >= s->n_kvmd) return -1;
    const sfaio_kvmd *k = &s->kvmd[set];
    if (first_frame) *first_frame = k->first_frame;
    if (pair < 0) return k->n_pairs;
    if (pair >= k->n_pairs) return -1;
    *key = k->keys[pair]; *val = k->vals[pair];
    return k->n_pairs;
}

/* ---------------- streaming repair ---------------- */

int sfaio_fixup_index(const char *path) {
    sfaio *s = sfaio_open(path);
    if (!s) return -1;
    FILE *fp = fopen(path, "r+b");
    if (!fp) { sfaio_close(s); return -1; }
    fseeko(fp, 0, SEEK_END);
    uint64_t fsize = (uint64_t)ftello(fp);
    uint64_t jm0;
    if (seek(fp, s->jtop_offset + 12 + 16) || rd(fp, &jm0, 8)) { fclose(fp); sfaio_close(s); return -1; }
    /* frames start after the first JMPF; later JMPF chunks are interleaved and skipped */
    uint64_t pos = jm0 + 12 + 8 + (uint64_t)s->jmpf_max * 32;
    uint32_t count = 0, l1 = 0;
    uint64_t jm = jm0;
    uint8_t *buf = malloc(s->frame_bytes);
    int codec = s->flags & 0xF;
    while (pos + 12 <= fsize) {
        char t[4]; uint64_t size;
        if (seek(fp, pos) || rd(fp, t, 4) || rd(fp, &size, 8) || size < 12 || pos + size > fsize) break;
        if (memcmp(t, "JMPF", 4) == 0) { pos += size; continue; }
        const int vec = memcmp(t, "FRVD", 4) == 0;
        if (!vec && memcmp(t, "FRMD", 4) != 0) break;
        /* FRMD: size counts the chunk header; FRVD: size = time(8) + payload, header not counted */
        uint64_t clen = vec ? size - 8 : size - 12, span = vec ? 12 + size : size;
        if (pos + span > fsize) break;
        double vt = 0;
        uint32_t crc = 0;                     /* vector frames are indexed without a CRC (as sfa.h does) */
        if (vec) {
            if (rd(fp, &vt, 8)) break;
        } else {
            uint8_t *comp = malloc(clen);
            if (rd(fp, comp, clen)) { free(comp); break; }
            int ok = decode_payload(s, codec, comp, clen, buf) == 0;
            free(comp);
            if (!ok) break;                   /* truncated or corrupt tail: stop here */
            crc = sfaio_crc32c(buf, s->frame_bytes);
        }
        uint32_t slot = count % s->jmpf_max;
        if (count > 0 && slot == 0) {         /* needs the next L2 table from JTOP */
            l1++;
            uint64_t nj = 0;
            if (seek(fp, s->jtop_offset + 12 + 16 + (uint64_t)l1 * 16) || rd(fp, &nj, 8) || nj == 0) break;
            jm = nj;
        }
        double tm = 0;
        uint32_t type = SFAIO_FRAME_VOXEL;
        uint64_t so = jm + 12 + 8 + (uint64_t)slot * 32;
        if (seek(fp, so) || rd(fp, &tm, 8)) break;
        if (vec) {                            /* the time is in the chunk; keep an I/P type already indexed */
            tm = vt;
            if (seek(fp, so + 28) || rd(fp, &type, 4)) break;
            if (type != SFAIO_FRAME_VEC_I && type != SFAIO_FRAME_VEC_P) type = SFAIO_FRAME_VEC_I;
        } else if (tm == 0.0 && count > 0) {
            tm = s->dt * count;
        }
        wr_at(fp, so, &tm, 8); wr(fp, &pos, 8); wr(fp, &clen, 8); wr(fp, &crc, 4); wr(fp, &type, 4);
        uint32_t inl2 = slot + 1;
        wr_at(fp, jm + 12 + 4, &inl2, 4);
        wr_at(fp, s->jtop_offset + 12 + 16 + (uint64_t)l1 * 16 + 12, &inl2, 4);
        count++;
        pos += span;
    }
    free(buf);
    uint32_t nl1 = l1 + 1;
    wr_at(fp, s->jtop_offset + 12 + 4, &nl1, 4);
    wr_at(fp, 68, &count, 4);
    uint32_t fl = s->flags & ~SFAIO_FLAG_STREAMING;
    wr_at(fp, 16, &fl, 4);
    fclose(fp);
    sfaio_close(s);
    return (int)count;
}
