// SuLi restart-file reader (host code; no device work).
//
// restart_salva (SuLi-IC/11_diversos.f90:581-599) writes the whole model
// state as ONE list-directed record, `write(6661,*) u, v, w, ku, kv, kw,
// prd0, prd1, prd, rho, ls_mu, ls, <boundary arrays>, ub, vb, wb, d_max,
// d_min, ls_m, hsx, hsy, hsz, hs, kurv, xmu, ymu, zmu, tf_u, tf_v, tf_w,
// vol_ini, vol_ins`, and restart_ini (:500-577) reads it back with the same
// list.  Each array is written whole in Fortran (column-major) order with
// the shape 0_modules.f90 allocates it with (:128-189, :273-277, :304-305,
// :334-337), so an array's values are found by counting.
//
// One array of the list, byxf (0_modules.f90:153), is declared but never
// allocated, so how many values a compiler writes for it is not fixed.  The
// reader therefore counts the whole record when it needs an array after
// byxf and takes byxf's count as what is left over.
//
// List-directed input rules followed here: values separated by blanks,
// line ends or one comma (with blanks around it); a null value (the target
// keeps its value) is "r*" (r of them), nothing between two commas, or a
// comma before the record's first value; "r*c" stands for r copies of c; a
// slash ends the record; reals may use E, D or Q exponents or a bare signed
// exponent ("1.5-100").  A value longer than the scanner's 127 characters
// is an error, not cut short.
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "suli_ps.h"

namespace {

// The record's arrays in write order, with their value counts.
struct Item {
    const char *name;
    long long count;   // -1: byxf, unknown
};

std::vector<Item> record_items(long long nx, long long ny, long long nz)
{
    const long long nu = (nx + 3) * (ny + 2) * (nz + 2), nv = (nx + 2) * (ny + 3) * (nz + 2),
                    nw = (nx + 2) * (ny + 2) * (nz + 3), nbox = (nx + 2) * (ny + 2) * (nz + 2), n = nx * ny * nz;
    const long long nfx = (nx + 1) * ny * nz, nfy = nx * (ny + 1) * nz, nfz = nx * ny * (nz + 1);
    // boundary arrays (0_modules.f90:176-186)
    const long long yz = (ny + 2) * (nz + 2), y1z = (ny + 3) * (nz + 2), yz1 = (ny + 2) * (nz + 3);
    const long long xz = (nx + 2) * (nz + 2), x1z = (nx + 3) * (nz + 2), xz1 = (nx + 2) * (nz + 3);
    const long long xy = (nx + 2) * (ny + 2), x1y = (nx + 3) * (ny + 2), xy1 = (nx + 2) * (ny + 3);
    return {
        {"u", nu},        {"v", nv},        {"w", nw},         {"ku", x1y},      {"kv", xy1},      {"kw", xy},
        {"prd0", nbox},   {"prd1", nbox},   {"prd", nbox},     {"rho", n},       {"ls_mu", n},     {"ls", n},
        {"bxx0", yz},     {"bxx1", yz},     {"bxy0", y1z},     {"bxz0", yz1},    {"bxxf", yz},     {"bxxf1", yz},
        {"bxyf", y1z},    {"bxzf", yz1},    {"byx0", x1z},     {"byy0", xz},     {"byy1", xz},     {"byz0", xz1},
        {"byxf", -1},     {"byyf", xz},     {"byyf1", xz},     {"byzf", xz1},    {"bzx0", x1y},    {"bzy0", xy1},
        {"bzz0", xy},     {"bzz1", xy},     {"bzxf", x1y},     {"bzyf", xy1},    {"bzzf", xy},     {"bzzf1", xy},
        {"ub", nu},       {"vb", nv},       {"wb", nw},        {"d_max", 1},     {"d_min", 1},     {"ls_m", 1},
        {"hsx", n},       {"hsy", n},       {"hsz", n},        {"hs", n},        {"kurv", n},      {"xmu", nfx},
        {"ymu", nfy},     {"zmu", nfz},     {"tf_u", nfx},     {"tf_v", nfy},    {"tf_w", nfz},    {"vol_ini", 1},
        {"vol_ins", 1},
    };
}

// Buffered scanner over the record's values.
class Scanner {
public:
    explicit Scanner(FILE *f) : f_(f), buf_(1 << 20) {}

    // Next value: returns 1 with *v set, 2 for a null value, 0 at the end
    // of the record, -1 on a malformed value (msg set).
    int next(double *v, std::string &msg)
    {
        if (rep_left_ > 0) {
            --rep_left_;
            if (rep_null_) return 2;
            *v = rep_val_;
            return 1;
        }
        if (ended_) return 0;
        char tok[128];
        int len = 0;
        int c;
        // skip the separator: blanks and line ends, and at most one comma; a
        // comma that finds one already taken (or the record's start) closes
        // a null value
        for (;;) {
            c = get();
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') continue;
            if (c != ',') break;
            if (comma_) return 2;
            comma_ = true;
        }
        if (c == EOF || c == '/') {
            ended_ = true;
            return 0;
        }
        while (c != EOF && !(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ',' || c == '/')) {
            if (len == (int)sizeof(tok) - 1) {
                msg = std::string("value longer than ") + std::to_string(sizeof(tok) - 1) + " characters: \"" +
                      std::string(tok, 24) + "...\"";
                return -1;
            }
            tok[len++] = (char)c;
            c = get();
        }
        // the value's own separator, when it is a comma
        comma_ = c == ',';
        if (c == '/') unget();
        tok[len] = 0;
        const char *star = std::strchr(tok, '*');
        if (star) {
            char *end = nullptr;
            const long long r = std::strtoll(tok, &end, 10);
            if (end != star || r < 1) {
                msg = std::string("bad repeat count in \"") + tok + "\"";
                return -1;
            }
            if (star[1] == 0) {
                rep_null_ = true;
            } else {
                rep_null_ = false;
                if (!parse_real(star + 1, &rep_val_)) {
                    msg = std::string("bad value \"") + tok + "\"";
                    return -1;
                }
            }
            rep_left_ = r;
            return next(v, msg);
        }
        if (!parse_real(tok, v)) {
            msg = std::string("bad value \"") + tok + "\"";
            return -1;
        }
        return 1;
    }

private:
    int get()
    {
        if (pos_ == have_) {
            have_ = std::fread(buf_.data(), 1, buf_.size(), f_);
            pos_ = 0;
            if (have_ == 0) return EOF;
        }
        return (unsigned char)buf_[pos_++];
    }
    void unget()
    {
        if (pos_ > 0) --pos_;
    }

    // Fortran real or integer constant: E/D/Q exponent letters, or a sign
    // directly after the mantissa standing for the exponent.
    static bool parse_real(const char *s, double *v)
    {
        char t[160];
        int o = 0;
        bool mant = false, expo = false;
        for (int i = 0; s[i]; ++i) {
            char ch = s[i];
            if (o > (int)sizeof(t) - 3) return false;
            if (ch == 'd' || ch == 'D' || ch == 'q' || ch == 'Q' || ch == 'e' || ch == 'E') {
                if (expo) return false;
                expo = true;
                t[o++] = 'e';
                continue;
            }
            if ((ch == '+' || ch == '-') && i > 0 && mant && !expo) {
                expo = true;
                t[o++] = 'e';
            }
            if ((ch >= '0' && ch <= '9') || ch == '.') mant = true;
            t[o++] = ch;
        }
        t[o] = 0;
        if (!mant) return false;
        char *end = nullptr;
        errno = 0;
        *v = std::strtod(t, &end);
        return end && *end == 0;
    }

    FILE *f_;
    std::vector<char> buf_;
    size_t pos_ = 0, have_ = 0;
    bool ended_ = false;
    bool comma_ = true;   // a comma closed the last value (the record's start counts as one)
    long long rep_left_ = 0;
    bool rep_null_ = false;
    double rep_val_ = 0.0;
};

int fail(char *err, size_t errlen, int code, const char *fmt, ...)
{
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

double *target_of(const suli_ps_restart *out, const std::string &name)
{
    if (name == "u") return out->u;
    if (name == "v") return out->v;
    if (name == "w") return out->w;
    if (name == "prd0") return out->prd0;
    if (name == "prd1") return out->prd1;
    if (name == "prd") return out->prd;
    if (name == "rho") return out->rho;
    if (name == "ls_mu") return out->ls_mu;
    if (name == "ls") return out->ls;
    if (name == "hs") return out->hs;
    return nullptr;
}

}  // namespace

extern "C" int suli_ps_restart_read(const char *path, int32_t nx, int32_t ny, int32_t nz, suli_ps_restart *out,
                                    char *err, size_t errlen)
{
    if (!path || !out) return fail(err, errlen, SULI_PS_EINVAL, "path and out must be non-NULL");
    if (nx < 1 || ny < 1 || nz < 1) return fail(err, errlen, SULI_PS_EINVAL, "grid %dx%dx%d", nx, ny, nz);
    const std::vector<Item> items = record_items(nx, ny, nz);
    // last item holding a requested array, and whether it lies past byxf
    int last = -1, byxf_at = -1;
    for (int i = 0; i < (int)items.size(); ++i) {
        if (items[i].count < 0) byxf_at = i;
        if (target_of(out, items[i].name)) last = i;
    }
    if (last < 0) return 0;
    long long byxf_count = 0;
    long long total = -1;
    if (last > byxf_at) {
        // count the whole record: byxf is what is left over
        FILE *f = std::fopen(path, "rb");
        if (!f) return fail(err, errlen, SULI_PS_EINVAL, "cannot open %s: %s", path, std::strerror(errno));
        Scanner sc(f);
        std::string msg;
        double v;
        int rc;
        total = 0;
        while ((rc = sc.next(&v, msg)) > 0) ++total;
        std::fclose(f);
        if (rc < 0) return fail(err, errlen, SULI_PS_EINVAL, "%s: %s", path, msg.c_str());
        long long known = 0;
        for (const Item &it : items)
            if (it.count > 0) known += it.count;
        byxf_count = total - known;
        if (byxf_count < 0)
            return fail(err, errlen, SULI_PS_EINVAL,
                        "%s holds %lld values, fewer than the %lld a %dx%dx%d restart record needs", path, total,
                        known, nx, ny, nz);
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(err, errlen, SULI_PS_EINVAL, "cannot open %s: %s", path, std::strerror(errno));
    Scanner sc(f);
    std::string msg;
    long long seen = 0;
    for (int i = 0; i <= last; ++i) {
        const long long cnt = items[i].count < 0 ? byxf_count : items[i].count;
        double *dst = target_of(out, items[i].name);
        for (long long k = 0; k < cnt; ++k) {
            double v;
            const int rc = sc.next(&v, msg);
            if (rc < 0) {
                std::fclose(f);
                return fail(err, errlen, SULI_PS_EINVAL, "%s: %s (value %lld, in %s)", path, msg.c_str(), seen,
                            items[i].name);
            }
            if (rc == 0) {
                std::fclose(f);
                return fail(err, errlen, SULI_PS_EINVAL, "%s ends after %lld values, inside %s (%lld of %lld)",
                            path, seen, items[i].name, k, cnt);
            }
            if (dst && rc == 1) dst[k] = v;
            ++seen;
        }
    }
    std::fclose(f);
    return 0;
}
