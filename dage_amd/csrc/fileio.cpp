// Binary function and grid files of DAGE (host code).
//
// .fun (Function3D_dump / Function3D_load, src/bubbles/function3d.F90:812-935;
// bubbles part Bubbles_write / Bubbles_read_and_init, src/bubbles/bubbles.F90:
// 712-756, 1251-1276): a Fortran unformatted stream file, little endian, no
// record markers:
//   byte  0  int64  cube section start (1-based position) or -1
//   byte  8  int64  pbc section start (written as zero, unused)
//   byte 16  int64  bubbles section start (1-based) or -1
//   byte 24  char[20] label (blank padded)
//   byte 44  int32  function type
//   cube section:    int64 shape[3], then the cube (x fastest) as real64
//   bubbles section: int32 nbub, nbub_global, lmax, k; per bubble real64 z,
//                    real64 centre[3], int32 global index (1-based); per
//                    global bubble real64 centre[3], real64 z; per bubble
//                    real64 f(npts, (lmax+1)^2)
// The radial grids are not in the file: the reader takes them from a mould.
//
// .g3d (Grid3D_dump / Grid3D_init_read, src/bubbles/grid.F90:1072-1145):
//   int32 nlip, int32 grid_type, then per axis int32 ncell, int32 nlip,
//   real64 qmin, real64 cell steps[ncell].
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dage_internal.hpp"

namespace dage {
namespace {

constexpr int kLabelLen = 20;

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

template <typename T>
bool put(FILE *f, const T &v) {
    return std::fwrite(&v, sizeof(T), 1, f) == 1;
}
bool put_n(FILE *f, const double *v, size_t n) { return std::fwrite(v, 8, n, f) == n; }
template <typename T>
bool get(FILE *f, T *v) {
    return std::fread(v, sizeof(T), 1, f) == 1;
}
bool get_n(FILE *f, double *v, size_t n) { return std::fread(v, 8, n, f) == n; }

}  // namespace
}  // namespace dage

using namespace dage;

extern "C" {

int dage_fun_write(const char *path, const char *label, int type, const int shape[3],
                   const double *cube, int nbub, int lmax, int k, const double *charges,
                   const double *centers, const int *npts, const double *const *bf) {
    DAGE_REQUIRE(path != nullptr, DAGE_ERR_ARG, "null path");
    DAGE_REQUIRE(!shape || cube, DAGE_ERR_ARG, "cube shape given without cube");
    DAGE_REQUIRE(nbub < 0 || (lmax >= 0 && (nbub == 0 || (charges && centers && npts && bf))),
                 DAGE_ERR_ARG, "incomplete bubbles");
    File fh;
    fh.f = std::fopen(path, "wb");
    DAGE_REQUIRE(fh.f != nullptr, DAGE_ERR_ARG, "cannot open %s for writing", path);
    FILE *f = fh.f;
    const int64_t zero = 0;
    int64_t cubestart = -1, bubblestart = -1;
    bool ok = put(f, cubestart) && put(f, zero) && put(f, bubblestart);
    char lab[kLabelLen];
    std::memset(lab, ' ', kLabelLen);
    if (label) std::memcpy(lab, label, std::min<size_t>(std::strlen(label), kLabelLen));
    ok = ok && std::fwrite(lab, 1, kLabelLen, f) == (size_t)kLabelLen;
    const int32_t t = type;
    ok = ok && put(f, t);
    if (shape) {
        cubestart = std::ftell(f) + 1;
        const size_t n = (size_t)shape[0] * shape[1] * shape[2];
        for (int a = 0; a < 3; ++a) ok = ok && put(f, (int64_t)shape[a]);
        ok = ok && put_n(f, cube, n);
    }
    if (nbub >= 0) {
        bubblestart = std::ftell(f) + 1;
        const int32_t nb = nbub, lm = lmax, kk = k;
        ok = ok && put(f, nb) && put(f, nb) && put(f, lm) && put(f, kk);
        for (int i = 0; i < nbub; ++i) {
            ok = ok && put(f, charges[i]) && put_n(f, centers + 3 * i, 3);
            ok = ok && put(f, (int32_t)(i + 1));
        }
        for (int i = 0; i < nbub; ++i) ok = ok && put_n(f, centers + 3 * i, 3) && put(f, charges[i]);
        const size_t nlm = (size_t)(lmax + 1) * (lmax + 1);
        for (int i = 0; i < nbub; ++i) ok = ok && put_n(f, bf[i], (size_t)npts[i] * nlm);
    }
    ok = ok && std::fseek(f, 0, SEEK_SET) == 0 && put(f, cubestart) &&
         std::fseek(f, 16, SEEK_SET) == 0 && put(f, bubblestart);
    DAGE_REQUIRE(ok, DAGE_ERR_ARG, "write to %s failed", path);
    return DAGE_OK;
}

int dage_fun_read_info(const char *path, char *label, int *type, int shape[3], int *nbub,
                       int *lmax, int *k) {
    DAGE_REQUIRE(path != nullptr, DAGE_ERR_ARG, "null path");
    File fh;
    fh.f = std::fopen(path, "rb");
    DAGE_REQUIRE(fh.f != nullptr, DAGE_ERR_ARG, "cannot open %s", path);
    FILE *f = fh.f;
    int64_t cubestart = 0, pbcstart = 0, bubblestart = 0;
    char lab[kLabelLen + 1] = {0};
    int32_t t = 0;
    bool ok = get(f, &cubestart) && get(f, &pbcstart) && get(f, &bubblestart) &&
              std::fread(lab, 1, kLabelLen, f) == (size_t)kLabelLen && get(f, &t);
    DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated header", path);
    if (label) std::memcpy(label, lab, kLabelLen + 1);
    if (type) *type = t;
    if (shape) {
        shape[0] = shape[1] = shape[2] = -1;
        if (cubestart != -1) {
            int64_t s[3];
            ok = std::fseek(f, (long)cubestart - 1, SEEK_SET) == 0 && get(f, &s[0]) &&
                 get(f, &s[1]) && get(f, &s[2]);
            DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated cube section", path);
            for (int a = 0; a < 3; ++a) shape[a] = (int)s[a];
        }
    }
    if (nbub) *nbub = -1;
    if (bubblestart != -1) {
        int32_t nb = 0, nbg = 0, lm = 0, kk = 0;
        ok = std::fseek(f, (long)bubblestart - 1, SEEK_SET) == 0 && get(f, &nb) && get(f, &nbg) &&
             get(f, &lm) && get(f, &kk);
        DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated bubbles section", path);
        if (nbub) *nbub = nb;
        if (lmax) *lmax = lm;
        if (k) *k = kk;
    }
    return DAGE_OK;
}

int dage_fun_read(const char *path, double *cube, double *charges, double *centers,
                  const int *npts, double *const *bf) {
    DAGE_REQUIRE(path != nullptr, DAGE_ERR_ARG, "null path");
    File fh;
    fh.f = std::fopen(path, "rb");
    DAGE_REQUIRE(fh.f != nullptr, DAGE_ERR_ARG, "cannot open %s", path);
    FILE *f = fh.f;
    int64_t cubestart = 0, pbcstart = 0, bubblestart = 0;
    bool ok = get(f, &cubestart) && get(f, &pbcstart) && get(f, &bubblestart);
    DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated header", path);
    if (cube && cubestart != -1) {
        int64_t s[3];
        ok = std::fseek(f, (long)cubestart - 1, SEEK_SET) == 0 && get(f, &s[0]) && get(f, &s[1]) &&
             get(f, &s[2]) && get_n(f, cube, (size_t)(s[0] * s[1] * s[2]));
        DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated cube", path);
    }
    if (bubblestart != -1 && (charges || centers || bf)) {
        int32_t nb = 0, nbg = 0, lm = 0, kk = 0;
        ok = std::fseek(f, (long)bubblestart - 1, SEEK_SET) == 0 && get(f, &nb) && get(f, &nbg) &&
             get(f, &lm) && get(f, &kk);
        for (int i = 0; ok && i < nb; ++i) {
            double z, c[3];
            int32_t idx;
            ok = get(f, &z) && get_n(f, c, 3) && get(f, &idx);
            if (charges) charges[i] = z;
            if (centers)
                for (int a = 0; a < 3; ++a) centers[3 * i + a] = c[a];
        }
        for (int i = 0; ok && i < nbg; ++i) {
            double c[3], z;
            ok = get_n(f, c, 3) && get(f, &z);
        }
        const size_t nlm = (size_t)(lm + 1) * (lm + 1);
        for (int i = 0; ok && i < nb; ++i) {
            DAGE_REQUIRE(!bf || npts, DAGE_ERR_ARG, "radial point counts needed");
            if (bf) ok = get_n(f, bf[i], (size_t)npts[i] * nlm);
        }
        DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated bubbles", path);
    }
    return DAGE_OK;
}

int dage_g3d_write(const char *path, int nlip, int grid_type, const int ncell[3],
                   const double qmin[3], const double *const steps[3]) {
    DAGE_REQUIRE(path && ncell && qmin && steps, DAGE_ERR_ARG, "null argument");
    File fh;
    fh.f = std::fopen(path, "wb");
    DAGE_REQUIRE(fh.f != nullptr, DAGE_ERR_ARG, "cannot open %s for writing", path);
    bool ok = put(fh.f, (int32_t)nlip) && put(fh.f, (int32_t)grid_type);
    for (int a = 0; a < 3; ++a)
        ok = ok && put(fh.f, (int32_t)ncell[a]) && put(fh.f, (int32_t)nlip) &&
             put(fh.f, qmin[a]) && put_n(fh.f, steps[a], ncell[a]);
    DAGE_REQUIRE(ok, DAGE_ERR_ARG, "write to %s failed", path);
    return DAGE_OK;
}

int dage_g3d_read(const char *path, int *nlip, int *grid_type, int ncell[3], double qmin[3],
                  double *const steps[3]) {
    DAGE_REQUIRE(path != nullptr, DAGE_ERR_ARG, "null path");
    File fh;
    fh.f = std::fopen(path, "rb");
    DAGE_REQUIRE(fh.f != nullptr, DAGE_ERR_ARG, "cannot open %s", path);
    int32_t nl = 0, gt = 0;
    bool ok = get(fh.f, &nl) && get(fh.f, &gt);
    if (nlip) *nlip = nl;
    if (grid_type) *grid_type = gt;
    for (int a = 0; ok && a < 3; ++a) {
        int32_t nc = 0, nla = 0;
        double q = 0.0;
        ok = get(fh.f, &nc) && get(fh.f, &nla) && get(fh.f, &q);
        DAGE_REQUIRE(!ok || nla == nl, DAGE_ERR_ARG,
                     "%s: each axis should have the same number of LIPs", path);
        if (ncell) ncell[a] = nc;
        if (qmin) qmin[a] = q;
        if (ok && steps && steps[a]) ok = get_n(fh.f, steps[a], nc);
        else if (ok) ok = std::fseek(fh.f, 8L * nc, SEEK_CUR) == 0;
    }
    DAGE_REQUIRE(ok, DAGE_ERR_ARG, "%s: truncated grid file", path);
    return DAGE_OK;
}

}  // extern "C"
