// sanitize_io_driver.cpp -- runs the instance-file reader and writer
// (fortkernels-_amd/csrc/fk_io.cpp, host code only) under AddressSanitizer
// and UndefinedBehaviorSanitizer.  Built and run by tests/test_io_sanitize.py
// (g++ -std=c++17 -fsanitize=address,undefined -Iinclude with
// fortkernels-_amd/csrc/fk_io.cpp and this file).
//
// fk_io.cpp reports errors through fk::set_error / fk::clear_error, which the
// library defines in another translation unit; this driver defines them as a
// plain message buffer.
//
// What it does, in a scratch directory given as argv[1]:
//   * writes a GPP problem (complex128 and complex64) and an FF problem, opens
//     each and compares every array with what was written;
//   * opens every prefix of each file (truncation) and copies with each
//     header byte replaced by 0x00, 0xff, 0x7f, 0x80 and its value with bit 0
//     or bit 6 flipped, plus targeted size and offset fields (INT32_MAX,
//     negative, unaligned, past the end).  A file the reader accepts must
//     satisfy an independent check of the documented layout (every array
//     inside the file, 64-byte aligned, after the header), and every byte of
//     every array it hands out is read;
//   * reads instance lists with comments, blank lines, CRLF ends, indented
//     and absolute paths, a 10000-character line and no final newline.
// Exits 1 on the first mismatch; sanitizer findings abort.
#include "fortkernels.h"

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace fk {
static char g_error[1024];
void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}
void clear_error() { g_error[0] = 0; }
}  // namespace fk

namespace {

int g_fail = 0;
uint64_t g_touched = 0;
std::string g_dir;

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s (last error: %s)\n", what, fk::g_error);
        std::exit(1);
    }
}

std::vector<char> slurp(const std::string &path)
{
    std::vector<char> v;
    FILE *f = std::fopen(path.c_str(), "rb");
    expect(f != nullptr, "open written file");
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) v.insert(v.end(), buf, buf + n);
    std::fclose(f);
    return v;
}

void spill(const std::string &path, const char *data, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    expect(f != nullptr, "create scratch file");
    if (n) expect(std::fwrite(data, 1, n, f) == n, "write scratch file");
    std::fclose(f);
}

// the layout's array extents from a header, overflow-checked; false if the
// header names an unknown kind or sizes the layout does not allow
bool layout(const std::vector<char> &file, std::vector<std::pair<int64_t, unsigned __int128>> &arrays)
{
    if (file.size() < 256) return false;
    int32_t kind, cb, sz[8];
    int64_t off[16];
    std::memcpy(&kind, file.data() + 8, 4);
    std::memcpy(&cb, file.data() + 12, 4);
    std::memcpy(sz, file.data() + 16, 32);
    std::memcpy(off, file.data() + 16 + 32 + 64, 128);
    using u128 = unsigned __int128;
    const u128 nb = (u128)(uint32_t)sz[0], C = (u128)(uint32_t)sz[2], G = (u128)(uint32_t)sz[3],
               nf = (u128)(uint32_t)sz[4], e = (u128)(uint32_t)cb;
    std::vector<u128> ext;
    if (kind == 1)
        ext = {8 * nf, 4 * G, 4 * (C + 1), 8 * C, e * nb * C, e * nb * C, e * G * C, e * G * C};
    else if (kind == 2)
        ext = {8 * nb, 8 * nf, 8 * nf, 16 * nf, 4 * G, 4 * (C + 1), 8 * C, 16 * nb * C, 16 * nb * C,
               16 * nf * G * C, 16 * nf * G * C};
    else
        return false;
    arrays.clear();
    for (size_t i = 0; i < ext.size(); ++i) arrays.push_back({off[i], ext[i]});
    return true;
}

uint64_t touch(const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    uint64_t s = 0;
    for (size_t i = 0; i < n; ++i) s += b[i];
    return s;
}

// Opens `path` (whose bytes are `file`); if the reader accepts it, checks
// the layout independently and reads every array it hands out.
bool try_open(const std::string &path, const std::vector<char> &file)
{
    fk_instance *inst = nullptr;
    const int rc = fk_instance_open(path.c_str(), &inst);
    if (rc != FK_OK) {
        expect(inst == nullptr, "rejected file leaves the handle NULL");
        expect(fk::g_error[0] != 0, "rejected file sets an error message");
        return false;
    }
    expect(inst != nullptr, "accepted file gives a handle");
    std::vector<std::pair<int64_t, unsigned __int128>> arr;
    expect(layout(file, arr), "accepted file has a known kind");
    for (const auto &a : arr) {
        expect(a.first >= 256 && a.first % 64 == 0, "accepted array offset is aligned and after the header");
        expect((unsigned __int128)a.first + a.second <= file.size(), "accepted array lies inside the file");
    }
    uint64_t s = 0;
    if (const fk_gpp_problem *p = fk_instance_gpp(inst)) {
        expect(fk_instance_kind(inst) == FK_INSTANCE_GPP && !fk_instance_ff(inst), "GPP accessors");
        const size_t e = (size_t)fk_instance_complex_bytes(inst), nb = p->number_bands, C = p->ncouls,
                     G = p->ngpown;
        expect(p->nvband >= 0 && p->nvband <= p->number_bands && C >= 1 && G >= 1 && p->nfreq >= 1,
               "accepted GPP sizes are in range");
        s += touch(p->wx, 8 * (size_t)p->nfreq) + touch(p->inv_igp_index, 4 * G) + touch(p->indinv, 4 * (C + 1)) +
             touch(p->vcoul, 8 * C) + touch(p->aqsmtemp, e * nb * C) + touch(p->aqsntemp, e * nb * C) +
             touch(p->I_eps_array, e * G * C) + touch(p->wtilde_array, e * G * C);
    } else {
        const fk_ff_full_problem *q = fk_instance_ff(inst);
        expect(q != nullptr && fk_instance_kind(inst) == FK_INSTANCE_FF, "FF accessors");
        const size_t nb = q->number_bands, C = q->ncouls, G = q->ngpown, nf = q->nfreq;
        expect(q->nvband >= 0 && q->nvband <= q->number_bands && nb >= 1 && C >= 1 && G >= 1 && nf >= 1 &&
                   q->nfreqeval >= 1,
               "accepted FF sizes are in range");
        s += touch(q->ekq, 8 * nb) + touch(q->dFreqGrid, 8 * nf) + touch(q->pref, 8 * nf) +
             touch(q->dFreqBrd, 16 * nf) + touch(q->inv_igp_index, 4 * G) + touch(q->indinv, 4 * (C + 1)) +
             touch(q->vcoul, 8 * C) + touch(q->aqsmtemp, 16 * nb * C) + touch(q->aqsntemp, 16 * nb * C) +
             touch(q->I_epsR_array, 16 * nf * G * C) + touch(q->I_epsA_array, 16 * nf * G * C);
    }
    g_touched += s;
    fk_instance_close(inst);
    return true;
}

// seeded bytes for the arrays
std::vector<char> fill(size_t n, unsigned seed)
{
    std::vector<char> v(n);
    uint32_t x = 2463534242u + seed;
    for (size_t i = 0; i < n; ++i) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
        v[i] = (char)(x >> 24);
    }
    return v;
}

void damage(const std::string &name, const std::string &path)
{
    const std::vector<char> good = slurp(path);
    const std::string scratch = g_dir + "/damaged.fki";
    int accepted = 0, total = 0;
    for (size_t len = 0; len < good.size(); ++len) {      // every prefix
        std::vector<char> f(good.begin(), good.begin() + (long)len);
        spill(scratch, f.data(), f.size());
        accepted += try_open(scratch, f);
        ++total;
    }
    const unsigned char vals[4] = {0x00, 0xff, 0x7f, 0x80};
    for (size_t i = 0; i < 256; ++i)                       // every header byte
        for (int k = 0; k < 6; ++k) {
            std::vector<char> f = good;
            const unsigned char o = (unsigned char)f[i];
            f[i] = (char)(k < 4 ? vals[k] : (k == 4 ? o ^ 0x01 : o ^ 0x40));
            if (f == good) continue;
            spill(scratch, f.data(), f.size());
            accepted += try_open(scratch, f);
            ++total;
        }
    // targeted fields: sizes[0..7] at 16, offsets[0..15] at 112
    const int32_t bad32[4] = {INT32_MAX, -1, INT32_MIN, 1 << 30};
    for (int s = 0; s < 8; ++s)
        for (int32_t v : bad32) {
            std::vector<char> f = good;
            std::memcpy(f.data() + 16 + 4 * s, &v, 4);
            spill(scratch, f.data(), f.size());
            accepted += try_open(scratch, f);
            ++total;
        }
    const int64_t bad64[6] = {0, 64, 257, (int64_t)good.size(), INT64_MAX, -64};
    for (int o = 0; o < 16; ++o)
        for (int64_t v : bad64) {
            std::vector<char> f = good;
            std::memcpy(f.data() + 112 + 8 * o, &v, 8);
            spill(scratch, f.data(), f.size());
            accepted += try_open(scratch, f);
            ++total;
        }
    std::printf("%s: %d damaged copies, %d accepted (each checked and read)\n", name.c_str(), total, accepted);
}

void round_trips()
{
    // GPP, nb 3, nv 1, ncouls 10, ngpown 4, nfreq 3
    const int nb = 3, C = 10, G = 4, nf = 3;
    for (int c64 = 0; c64 <= 1; ++c64) {
        const size_t e = c64 ? 8 : 16;
        std::vector<char> wx = fill(8 * nf, 1), vc = fill(8 * C, 2), am = fill(e * nb * C, 3), an = fill(e * nb * C, 4),
                          ie = fill(e * G * C, 5), wt = fill(e * G * C, 6);
        std::vector<int32_t> inv = {0, 2, 5, 10}, ind(C + 1);
        for (int i = 0; i <= C; ++i) ind[i] = i < C ? i : C - 1;
        fk_gpp_problem p;
        std::memset(&p, 0, sizeof p);
        p.number_bands = nb; p.nvband = 1; p.ncouls = C; p.ngpown = G; p.nfreq = nf;
        p.wx = reinterpret_cast<const double *>(wx.data());
        p.aqsmtemp = am.data(); p.aqsntemp = an.data(); p.I_eps_array = ie.data(); p.wtilde_array = wt.data();
        p.vcoul = reinterpret_cast<const double *>(vc.data());
        p.inv_igp_index = inv.data(); p.indinv = ind.data();
        p.to1 = 1e-6; p.limittwo = 0.25; p.limitone = 2.5e5; p.occ = 1.0;
        const std::string path = g_dir + (c64 ? "/gpp64.fki" : "/gpp128.fki");
        expect(fk_gpp_instance_write(path.c_str(), &p, c64) == FK_OK, "GPP write");
        fk_instance *inst = nullptr;
        expect(fk_instance_open(path.c_str(), &inst) == FK_OK, "GPP open");
        const fk_gpp_problem *q = fk_instance_gpp(inst);
        expect(q && q->number_bands == nb && q->nvband == 1 && q->ncouls == C && q->ngpown == G && q->nfreq == nf,
               "GPP sizes");
        expect(q->to1 == p.to1 && q->limittwo == p.limittwo && q->limitone == p.limitone && q->occ == p.occ,
               "GPP scalars");
        expect(!std::memcmp(q->wx, wx.data(), wx.size()) && !std::memcmp(q->vcoul, vc.data(), vc.size()) &&
                   !std::memcmp(q->aqsmtemp, am.data(), am.size()) && !std::memcmp(q->aqsntemp, an.data(), an.size()) &&
                   !std::memcmp(q->I_eps_array, ie.data(), ie.size()) &&
                   !std::memcmp(q->wtilde_array, wt.data(), wt.size()) &&
                   !std::memcmp(q->inv_igp_index, inv.data(), 4 * G) && !std::memcmp(q->indinv, ind.data(), 4 * (C + 1)),
               "GPP arrays");
        expect(fk_instance_complex_bytes(inst) == (int32_t)e, "GPP element size");
        fk_instance_close(inst);
        damage(c64 ? "GPP complex64" : "GPP complex128", path);
    }
    // FF, nb 3, nv 1, ncouls 6, ngpown 2, nfreq 4, nfreqeval 5
    {
        const int fnb = 3, fC = 6, fG = 2, fnf = 4;
        std::vector<char> ekq = fill(8 * fnb, 11), grid = fill(8 * fnf, 12), pref = fill(8 * fnf, 13),
                          brd = fill(16 * fnf, 14), vc = fill(8 * fC, 15), am = fill(16 * fnb * fC, 16),
                          an = fill(16 * fnb * fC, 17), ir = fill(16 * fnf * fG * fC, 18),
                          ia = fill(16 * fnf * fG * fC, 19);
        std::vector<int32_t> inv = {1, 4}, ind(fC + 1);
        for (int i = 0; i <= fC; ++i) ind[i] = i < fC ? i : fC - 1;
        fk_ff_full_problem p;
        std::memset(&p, 0, sizeof p);
        p.number_bands = fnb; p.nvband = 1; p.ncouls = fC; p.ngpown = fG; p.nfreq = fnf; p.nfreqeval = 5;
        p.aqsmtemp = am.data(); p.aqsntemp = an.data(); p.I_epsR_array = ir.data(); p.I_epsA_array = ia.data();
        p.vcoul = reinterpret_cast<const double *>(vc.data());
        p.inv_igp_index = inv.data(); p.indinv = ind.data();
        p.ekq = reinterpret_cast<const double *>(ekq.data());
        p.dFreqGrid = reinterpret_cast<const double *>(grid.data());
        p.pref = reinterpret_cast<const double *>(pref.data());
        p.dFreqBrd = reinterpret_cast<const fk_cdouble *>(brd.data());
        p.pref_zb = 0.5 / 3.14; p.freqevalmin = 0.0; p.freqevalstep = 0.5; p.occ = 1.0; p.achs_freq = 2;
        const std::string path = g_dir + "/ff.fki";
        expect(fk_ff_instance_write(path.c_str(), &p) == FK_OK, "FF write");
        fk_instance *inst = nullptr;
        expect(fk_instance_open(path.c_str(), &inst) == FK_OK, "FF open");
        const fk_ff_full_problem *q = fk_instance_ff(inst);
        expect(q && !fk_instance_gpp(inst) && q->number_bands == fnb && q->nvband == 1 && q->ncouls == fC &&
                   q->ngpown == fG && q->nfreq == fnf && q->nfreqeval == 5 && q->achs_freq == 2,
               "FF sizes");
        expect(!std::memcmp(q->ekq, ekq.data(), ekq.size()) && !std::memcmp(q->dFreqGrid, grid.data(), grid.size()) &&
                   !std::memcmp(q->pref, pref.data(), pref.size()) &&
                   !std::memcmp(q->dFreqBrd, brd.data(), brd.size()) && !std::memcmp(q->vcoul, vc.data(), vc.size()) &&
                   !std::memcmp(q->aqsmtemp, am.data(), am.size()) && !std::memcmp(q->aqsntemp, an.data(), an.size()) &&
                   !std::memcmp(q->I_epsR_array, ir.data(), ir.size()) &&
                   !std::memcmp(q->I_epsA_array, ia.data(), ia.size()) &&
                   !std::memcmp(q->inv_igp_index, inv.data(), 4 * fG) &&
                   !std::memcmp(q->indinv, ind.data(), 4 * (fC + 1)),
               "FF arrays");
        fk_instance_close(inst);
        damage("FF", path);
    }
    // writer argument checks
    fk_gpp_problem z;
    std::memset(&z, 0, sizeof z);
    expect(fk_gpp_instance_write((g_dir + "/x.fki").c_str(), &z, 0) == FK_ERR_ARG, "GPP write rejects zero sizes");
    expect(fk_gpp_instance_write(nullptr, &z, 0) == FK_ERR_ARG, "GPP write rejects a NULL path");
    fk_ff_full_problem zf;
    std::memset(&zf, 0, sizeof zf);
    expect(fk_ff_instance_write((g_dir + "/x.fki").c_str(), &zf) == FK_ERR_ARG, "FF write rejects zero sizes");
    fk_instance *inst = nullptr;
    expect(fk_instance_open((g_dir + "/does-not-exist.fki").c_str(), &inst) == FK_ERR_ARG && !inst,
           "open rejects a missing file");
    expect(fk_instance_open(nullptr, &inst) == FK_ERR_ARG, "open rejects a NULL path");
    expect(fk_instance_kind(nullptr) == 0 && !fk_instance_gpp(nullptr) && !fk_instance_ff(nullptr) &&
               fk_instance_complex_bytes(nullptr) == 0,
           "accessors accept NULL");
    fk_instance_close(nullptr);
}

void lists()
{
    const std::string path = g_dir + "/list.txt";
    std::string longname(10000, 'a');
    const std::string text = "# a comment\n\n   \t\nfirst.fki\r\n  second.fki  \n/abs/third.fki\n" + longname +
                             "\n#tail comment\nlast.fki";
    spill(path, text.data(), text.size());
    char **paths = nullptr;
    int32_t n = -1;
    expect(fk_instance_list_read(path.c_str(), &paths, &n) == FK_OK, "list read");
    expect(n == 5, "list entries");
    expect(std::string(paths[0]) == g_dir + "/first.fki", "list relative path");
    expect(std::string(paths[1]) == g_dir + "/second.fki", "list trimmed path");
    expect(std::string(paths[2]) == "/abs/third.fki", "list absolute path");
    expect(std::string(paths[3]) == g_dir + "/" + longname, "list long line kept whole");
    expect(std::string(paths[4]) == g_dir + "/last.fki", "list last line without newline");
    expect(paths[5] == nullptr, "list is NULL-terminated");
    fk_instance_list_free(paths, n);
    spill(path, "", 0);
    expect(fk_instance_list_read(path.c_str(), &paths, &n) == FK_OK && n == 0, "empty list");
    fk_instance_list_free(paths, n);
    expect(fk_instance_list_read((g_dir + "/missing.txt").c_str(), &paths, &n) == FK_ERR_ARG && n == 0 && !paths,
           "missing list");
    fk_instance_list_free(nullptr, 0);
    std::printf("instance lists: ok\n");
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s SCRATCH_DIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    round_trips();
    lists();
    std::printf("sanitize_io_driver: all cases passed (%llu)\n", (unsigned long long)(g_touched & 0xffff));
    return g_fail ? 1 : 0;
}
