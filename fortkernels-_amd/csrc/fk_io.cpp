// fk_io.cpp -- instance files: one GPP or FF problem (sizes, constants and
// every input array in the C layouts of the ABI) in a single binary file,
// for inputs dumped from a real BerkeleyGW run, and the instance lists of
// batched runs.  The reference has no file format (its programs synthesise
// constant inputs, BGW-Kernels/gppKer.cpp:274-291, FFKer_cori/ffKernel.cpp:
// 395-450); the layouts are the reference's arrays, so a Fortran dump of
// aqsmtemp(ncouls, number_bands), I_eps_array(ncouls, ngpown), ... is already
// in this order.
//
// Layout (little-endian):
//   [0, 256)  header: char magic[8] = "FKINST01"; int32 kind (1 GPP, 2 FF),
//             int32 complex_bytes (16 complex128, 8 complex64 (GPP only));
//             int32 sizes[8]; double scalars[8]; int64 offsets[16]; zero pad
//   arrays    each at its offset, 64-byte aligned, back to back
// GPP sizes:   number_bands, nvband, ncouls, ngpown, nfreq
//     scalars: to1, limittwo, limitone, occ
//     arrays:  0 wx[nfreq] f64, 1 inv_igp_index[ngpown] i32,
//              2 indinv[ncouls+1] i32, 3 vcoul[ncouls] f64,
//              4 aqsmtemp[nb][ncouls], 5 aqsntemp[nb][ncouls],
//              6 I_eps_array[ngpown][ncouls], 7 wtilde_array[ngpown][ncouls]
// FF sizes:    number_bands, nvband, ncouls, ngpown, nfreq, nfreqeval, achs_freq
//     scalars: pref_zb, freqevalmin, freqevalstep, occ
//     arrays:  0 ekq[nb] f64, 1 dFreqGrid[nfreq] f64, 2 pref[nfreq] f64,
//              3 dFreqBrd[nfreq] c128, 4 inv_igp_index[ngpown] i32,
//              5 indinv[ncouls+1] i32, 6 vcoul[ncouls] f64,
//              7 aqsmtemp[nb][ncouls], 8 aqsntemp[nb][ncouls],
//              9 I_epsR[nfreq][ngpown][ncouls], 10 I_epsA[...] (c128)
// A reader maps the file read-only; the problem it hands out points into
// the mapping (host memory: FK_MEM_HOST, or copy to the device).
#include "../../include/fortkernels.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace fk {
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
}  // namespace fk

using fk::set_error;

namespace {

constexpr char kMagic[8] = {'F', 'K', 'I', 'N', 'S', 'T', '0', '1'};
constexpr size_t kHeader = 256, kAlign = 64;

struct Header {
    char magic[8];
    int32_t kind, complex_bytes;
    int32_t sizes[8];
    double scalars[8];
    int64_t offsets[16];
};
static_assert(sizeof(Header) <= kHeader, "header fits its 256 bytes");

struct Piece {
    const void *data;
    size_t bytes;
};

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

int write_file(const char *path, Header &h, const std::vector<Piece> &pieces)
{
    size_t off = kHeader;
    for (size_t i = 0; i < pieces.size(); ++i) {
        h.offsets[i] = (int64_t)off;
        off = align_up(off + pieces[i].bytes);
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) { set_error("instance write: cannot open %s: %s", path, std::strerror(errno)); return FK_ERR_ARG; }
    char head[kHeader] = {};
    std::memcpy(head, &h, sizeof(h));
    bool ok = std::fwrite(head, 1, kHeader, f) == kHeader;
    static const char zeros[kAlign] = {};
    size_t pos = kHeader;
    for (size_t i = 0; ok && i < pieces.size(); ++i) {
        if (pieces[i].bytes) ok = std::fwrite(pieces[i].data, 1, pieces[i].bytes, f) == pieces[i].bytes;
        pos += pieces[i].bytes;
        const size_t pad = align_up(pos) - pos;
        if (ok && pad) ok = std::fwrite(zeros, 1, pad, f) == pad;
        pos += pad;
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) { set_error("instance write: short write to %s", path); return FK_ERR_ARG; }
    return FK_OK;
}

}  // namespace

struct fk_instance {
    void *map = nullptr;
    size_t bytes = 0;
    int32_t kind = 0, complex_bytes = 0;
    fk_gpp_problem gpp{};
    fk_ff_full_problem ff{};     // semantics 0: an instance holds inputs, not a choice of evaluation
};

extern "C" {

int fk_gpp_instance_write(const char *path, const fk_gpp_problem *p, int32_t complex64)
{
    fk::clear_error();
    if (!path || !p || p->number_bands < 0 || p->nvband < 0 || p->nvband > p->number_bands || p->ncouls < 1 ||
        p->ngpown < 1 || p->nfreq < 1 || !p->wx || !p->inv_igp_index || !p->indinv || !p->vcoul ||
        (p->number_bands > 0 && (!p->aqsmtemp || !p->aqsntemp)) || !p->I_eps_array || !p->wtilde_array) {
        set_error("fk_gpp_instance_write: NULL argument or bad sizes");
        return FK_ERR_ARG;
    }
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.kind = 1;
    h.complex_bytes = complex64 ? 8 : 16;
    const int32_t sz[5] = {p->number_bands, p->nvband, p->ncouls, p->ngpown, p->nfreq};
    std::memcpy(h.sizes, sz, sizeof(sz));
    h.scalars[0] = p->to1; h.scalars[1] = p->limittwo; h.scalars[2] = p->limitone; h.scalars[3] = p->occ;
    const size_t e = (size_t)h.complex_bytes, nbC = (size_t)p->number_bands * p->ncouls,
                 GC = (size_t)p->ngpown * p->ncouls;
    std::vector<Piece> pc = {{p->wx, sizeof(double) * p->nfreq},
                             {p->inv_igp_index, sizeof(int32_t) * p->ngpown},
                             {p->indinv, sizeof(int32_t) * ((size_t)p->ncouls + 1)},
                             {p->vcoul, sizeof(double) * p->ncouls},
                             {p->aqsmtemp, e * nbC},
                             {p->aqsntemp, e * nbC},
                             {p->I_eps_array, e * GC},
                             {p->wtilde_array, e * GC}};
    return write_file(path, h, pc);
}

int fk_ff_instance_write(const char *path, const fk_ff_full_problem *p)
{
    fk::clear_error();
    if (!path || !p || p->number_bands < 1 || p->nvband < 0 || p->nvband > p->number_bands || p->ncouls < 1 ||
        p->ngpown < 1 || p->nfreq < 1 || p->nfreqeval < 1 || !p->ekq || !p->dFreqGrid || !p->pref ||
        !p->dFreqBrd || !p->inv_igp_index || !p->indinv || !p->vcoul || !p->aqsmtemp || !p->aqsntemp ||
        !p->I_epsR_array || !p->I_epsA_array) {
        set_error("fk_ff_instance_write: NULL argument or bad sizes");
        return FK_ERR_ARG;
    }
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.kind = 2;
    h.complex_bytes = 16;
    const int32_t sz[7] = {p->number_bands, p->nvband, p->ncouls, p->ngpown, p->nfreq, p->nfreqeval, p->achs_freq};
    std::memcpy(h.sizes, sz, sizeof(sz));
    h.scalars[0] = p->pref_zb; h.scalars[1] = p->freqevalmin; h.scalars[2] = p->freqevalstep; h.scalars[3] = p->occ;
    const size_t nb = (size_t)p->number_bands, nf = (size_t)p->nfreq, C = (size_t)p->ncouls, G = (size_t)p->ngpown;
    std::vector<Piece> pc = {{p->ekq, sizeof(double) * nb},
                             {p->dFreqGrid, sizeof(double) * nf},
                             {p->pref, sizeof(double) * nf},
                             {p->dFreqBrd, sizeof(fk_cdouble) * nf},
                             {p->inv_igp_index, sizeof(int32_t) * G},
                             {p->indinv, sizeof(int32_t) * (C + 1)},
                             {p->vcoul, sizeof(double) * C},
                             {p->aqsmtemp, sizeof(fk_cdouble) * nb * C},
                             {p->aqsntemp, sizeof(fk_cdouble) * nb * C},
                             {p->I_epsR_array, sizeof(fk_cdouble) * nf * G * C},
                             {p->I_epsA_array, sizeof(fk_cdouble) * nf * G * C}};
    return write_file(path, h, pc);
}

int fk_instance_open(const char *path, fk_instance **out)
{
    fk::clear_error();
    if (!path || !out) { set_error("fk_instance_open: NULL argument"); return FK_ERR_ARG; }
    *out = nullptr;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) { set_error("fk_instance_open: cannot open %s: %s", path, std::strerror(errno)); return FK_ERR_ARG; }
    struct stat st;
    if (::fstat(fd, &st) != 0 || st.st_size < (off_t)kHeader) {
        ::close(fd);
        set_error("fk_instance_open: %s is not an instance file (too short)", path);
        return FK_ERR_ARG;
    }
    const size_t bytes = (size_t)st.st_size;
    void *map = ::mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) { set_error("fk_instance_open: mmap %s: %s", path, std::strerror(errno)); return FK_ERR_ARG; }
    Header h;
    std::memcpy(&h, map, sizeof(h));
    const char *base = static_cast<const char *>(map);
    auto fail = [&](const char *why) {
        ::munmap(map, bytes);
        set_error("fk_instance_open: %s: %s", path, why);
        return FK_ERR_ARG;
    };
    if (std::memcmp(h.magic, kMagic, 8) != 0) return fail("bad magic (not an FKINST01 file)");
    if (h.kind != 1 && h.kind != 2) return fail("unknown instance kind");
    if (h.complex_bytes != 16 && !(h.kind == 1 && h.complex_bytes == 8)) return fail("bad complex element size");
    for (int i = 0; i < 8; ++i)
        if (h.sizes[i] < 0) return fail("negative size");
    // expected extents of the arrays, in file order
    std::vector<size_t> ext;
    const size_t e = (size_t)h.complex_bytes;
    // a * b * c bytes, or SIZE_MAX when the product does not fit (then no
    // file is large enough and the extent check below rejects it)
    auto prod = [](size_t a, size_t b, size_t c) -> size_t {
        size_t ab, abc;
        if (__builtin_mul_overflow(a, b, &ab) || __builtin_mul_overflow(ab, c, &abc)) return SIZE_MAX;
        return abc;
    };
    if (h.kind == 1) {
        const size_t nb = h.sizes[0], C = h.sizes[2], G = h.sizes[3], nf = h.sizes[4];
        if (C < 1 || G < 1 || nf < 1 || h.sizes[1] > h.sizes[0]) return fail("bad GPP sizes");
        ext = {8 * nf, 4 * G, 4 * (C + 1), 8 * C, prod(e, nb, C), prod(e, nb, C), prod(e, G, C), prod(e, G, C)};
    } else {
        const size_t nb = h.sizes[0], C = h.sizes[2], G = h.sizes[3], nf = h.sizes[4];
        if (nb < 1 || C < 1 || G < 1 || nf < 1 || h.sizes[5] < 1 || h.sizes[1] > h.sizes[0]) return fail("bad FF sizes");
        const size_t fgc = prod(nf, G, C);
        ext = {8 * nb, 8 * nf, 8 * nf, 16 * nf, 4 * G, 4 * (C + 1), 8 * C, prod(16, nb, C), prod(16, nb, C),
               prod(16, fgc, 1), prod(16, fgc, 1)};
    }
    std::vector<const void *> ptr(ext.size());
    for (size_t i = 0; i < ext.size(); ++i) {
        const int64_t off = h.offsets[i];
        if (off < (int64_t)kHeader || off % (int64_t)kAlign != 0 || (size_t)off > bytes || ext[i] > bytes - (size_t)off)
            return fail("array offset or extent outside the file");
        ptr[i] = base + off;
    }
    fk_instance *inst = new (std::nothrow) fk_instance;
    if (!inst) { ::munmap(map, bytes); set_error("fk_instance_open: out of memory"); return FK_ERR_NOMEM; }
    inst->map = map;
    inst->bytes = bytes;
    inst->kind = h.kind;
    inst->complex_bytes = h.complex_bytes;
    std::memset(&inst->gpp, 0, sizeof(inst->gpp));
    std::memset(&inst->ff, 0, sizeof(inst->ff));
    if (h.kind == 1) {
        fk_gpp_problem &p = inst->gpp;
        p.number_bands = h.sizes[0]; p.nvband = h.sizes[1]; p.ncouls = h.sizes[2]; p.ngpown = h.sizes[3];
        p.nfreq = h.sizes[4];
        p.to1 = h.scalars[0]; p.limittwo = h.scalars[1]; p.limitone = h.scalars[2]; p.occ = h.scalars[3];
        p.wx = static_cast<const double *>(ptr[0]);
        p.inv_igp_index = static_cast<const int32_t *>(ptr[1]);
        p.indinv = static_cast<const int32_t *>(ptr[2]);
        p.vcoul = static_cast<const double *>(ptr[3]);
        p.aqsmtemp = ptr[4]; p.aqsntemp = ptr[5]; p.I_eps_array = ptr[6]; p.wtilde_array = ptr[7];
    } else {
        fk_ff_full_problem &p = inst->ff;
        p.number_bands = h.sizes[0]; p.nvband = h.sizes[1]; p.ncouls = h.sizes[2]; p.ngpown = h.sizes[3];
        p.nfreq = h.sizes[4]; p.nfreqeval = h.sizes[5]; p.achs_freq = h.sizes[6];
        p.pref_zb = h.scalars[0]; p.freqevalmin = h.scalars[1]; p.freqevalstep = h.scalars[2]; p.occ = h.scalars[3];
        p.ekq = static_cast<const double *>(ptr[0]);
        p.dFreqGrid = static_cast<const double *>(ptr[1]);
        p.pref = static_cast<const double *>(ptr[2]);
        p.dFreqBrd = static_cast<const fk_cdouble *>(ptr[3]);
        p.inv_igp_index = static_cast<const int32_t *>(ptr[4]);
        p.indinv = static_cast<const int32_t *>(ptr[5]);
        p.vcoul = static_cast<const double *>(ptr[6]);
        p.aqsmtemp = ptr[7]; p.aqsntemp = ptr[8]; p.I_epsR_array = ptr[9]; p.I_epsA_array = ptr[10];
    }
    *out = inst;
    return FK_OK;
}

int fk_instance_kind(const fk_instance *inst) { return inst ? inst->kind : 0; }

int32_t fk_instance_complex_bytes(const fk_instance *inst) { return inst ? inst->complex_bytes : 0; }

const fk_gpp_problem *fk_instance_gpp(const fk_instance *inst)
{
    return inst && inst->kind == FK_INSTANCE_GPP ? &inst->gpp : nullptr;
}

const fk_ff_full_problem *fk_instance_ff(const fk_instance *inst)
{
    return inst && inst->kind == FK_INSTANCE_FF ? &inst->ff : nullptr;
}

void fk_instance_close(fk_instance *inst)
{
    if (!inst) return;
    if (inst->map) ::munmap(inst->map, inst->bytes);
    delete inst;
}

int fk_instance_list_read(const char *path, char ***paths, int32_t *count)
{
    fk::clear_error();
    if (!path || !paths || !count) { set_error("fk_instance_list_read: NULL argument"); return FK_ERR_ARG; }
    *paths = nullptr;
    *count = 0;
    FILE *f = std::fopen(path, "r");
    if (!f) { set_error("fk_instance_list_read: cannot open %s: %s", path, std::strerror(errno)); return FK_ERR_ARG; }
    std::string dir(path);
    const size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);
    std::vector<std::string> items;
    char *line = nullptr;       // getline grows it: a path of any length stays one entry
    size_t cap = 0;
    while (::getline(&line, &cap, f) >= 0) {
        std::string s(line);
        while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) s.pop_back();
        size_t b = 0;
        while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) ++b;
        s = s.substr(b);
        if (s.empty() || s[0] == '#') continue;
        items.push_back(s[0] == '/' ? s : dir + s);
    }
    std::free(line);
    std::fclose(f);
    char **arr = static_cast<char **>(std::calloc(items.size() + 1, sizeof(char *)));
    if (!arr) { set_error("fk_instance_list_read: out of memory"); return FK_ERR_NOMEM; }
    for (size_t i = 0; i < items.size(); ++i) {
        arr[i] = static_cast<char *>(std::malloc(items[i].size() + 1));
        if (!arr[i]) {
            fk_instance_list_free(arr, (int32_t)i);
            set_error("fk_instance_list_read: out of memory");
            return FK_ERR_NOMEM;
        }
        std::memcpy(arr[i], items[i].c_str(), items[i].size() + 1);
    }
    *paths = arr;
    *count = (int32_t)items.size();
    return FK_OK;
}

void fk_instance_list_free(char **paths, int32_t count)
{
    if (!paths) return;
    for (int32_t i = 0; i < count; ++i) std::free(paths[i]);
    std::free(paths);
}

}  // extern "C"
