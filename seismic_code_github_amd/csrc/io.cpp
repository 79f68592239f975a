// Files on either side of the propagators: the Fortran drivers' parameter
// cards (.par) and their SU shot gathers.  Host code only.
//
//   Q_forward read_par            Q sub:121-234      -> sfd_read_q_par
//   yang read_par                 pub:48-145         -> sfd_read_iso_par
//   write_currshot_disk           Q sub:1476-1546,
//                                 pub:1533-1590      -> sfd_write_su_shot
//   merge_shot_files              Q sub:1550-1624,
//                                 pub:1594-1684      -> sfd_merge_su
//   type segy (240-byte header)   module:32-116
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "seisfd.h"

namespace {

// The card's lines in order; the reader alternates a comment line and a
// value line as the Fortran reads do.
struct Card {
    std::vector<std::string> lines;
    size_t next = 0;
    std::string err;

    bool open(const char* path) {
        std::ifstream in(path);
        if (!in) {
            err = std::string("cannot read ") + path;
            return false;
        }
        std::string s;
        while (std::getline(in, s)) {
            if (!s.empty() && s.back() == '\r') s.pop_back();
            lines.push_back(s);
        }
        return true;
    }
    // read(10,'(A)'): the whole next line.
    bool line(std::string* out) {
        if (next >= lines.size()) {
            err = "card ends early (line " + std::to_string(next + 1) + ")";
            return false;
        }
        *out = lines[next++];
        return true;
    }
    // A comment line then a value line.
    bool value_line(std::string* out) {
        std::string skip;
        return line(&skip) && line(out);
    }
};

// Trailing blanks removed, then leading ones (Fortran trim(adjustl(.))).
std::string trimmed(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t");
    return s.substr(a, b - a + 1);
}

// List-directed read (read(10,*)) of n numbers from one line: values
// separated by blanks or commas; '/' ends the list.
bool list_reals(const std::string& s, int n, double* v, std::string* err) {
    std::string t = s;
    const size_t slash = t.find('/');
    if (slash != std::string::npos) t.resize(slash);
    for (char& c : t)
        if (c == ',' || c == '\t') c = ' ';
    const char* p = t.c_str();
    for (int q = 0; q < n; ++q) {
        while (*p == ' ') ++p;
        if (!*p) {
            *err = "expected " + std::to_string(n) + " values in '" + s + "'";
            return false;
        }
        char* e = nullptr;
        // Fortran also writes exponents with d / D.
        std::string tok;
        while (*p && *p != ' ') tok += (*p == 'd' || *p == 'D') ? 'e' : *p, ++p;
        v[q] = std::strtod(tok.c_str(), &e);
        if (!e || *e) {
            *err = "not a number: '" + tok + "'";
            return false;
        }
    }
    return true;
}

// Intel Fortran's width-less edit descriptors on input: '(F)' is F15.7 and
// '(I)' I12 — only the first 15 (12) characters count, blanks inside the
// field are ignored, and an F field without a decimal point is scaled by
// 10^-7.
bool f_field(const std::string& s, double* v, std::string* err) {
    std::string f;
    for (size_t i = 0; i < s.size() && i < 15; ++i)
        if (s[i] != ' ' && s[i] != '\t') f += (s[i] == 'd' || s[i] == 'D') ? 'e' : s[i];
    if (f.empty()) {
        *v = 0.0;  // an all-blank field reads as zero
        return true;
    }
    char* e = nullptr;
    double x = std::strtod(f.c_str(), &e);
    if (!e || *e) {
        *err = "not an F15.7 field: '" + s + "'";
        return false;
    }
    if (f.find('.') == std::string::npos) {
        // Implied decimal point: d = 7 digits after it (in front of any exponent).
        x *= 1e-7;
    }
    *v = x;
    return true;
}

bool i_field(const std::string& s, int* v, std::string* err) {
    std::string f;
    for (size_t i = 0; i < s.size() && i < 12; ++i)
        if (s[i] != ' ' && s[i] != '\t') f += s[i];
    if (f.empty()) {
        *v = 0;
        return true;
    }
    char* e = nullptr;
    const long x = std::strtol(f.c_str(), &e, 10);
    if (!e || *e) {
        *err = "not an I12 field: '" + s + "'";
        return false;
    }
    *v = (int)x;
    return true;
}

void copy_name(char* dst, const std::string& s) {
    std::strncpy(dst, trimmed(s).c_str(), 255);
    dst[255] = 0;
}

int card_fail(char* err, const std::string& what) {
    if (err) {
        std::strncpy(err, what.c_str(), 255);
        err[255] = 0;
    }
    return SFD_EINVAL;
}

// The 240-byte SU trace header (module:32-116), the fields the writers set.
#pragma pack(push, 1)
struct SuHeader {
    int32_t tracl, tracr, fldr, tracf, ep, cdp, cdpt;
    int16_t trid, nvs, nhs, duse;
    int32_t offset, gelev, selev, sdepth, gdel, sdel, swdep, gwdep;
    int16_t scalel, scalco;
    int32_t sx, sy, gx, gy;
    int16_t counit, wevel, swevel, sut, gut, sstat, gstat, tstat, laga, lagb, delrt, muts, mute;
    uint16_t ns, dt;
    int16_t rest[31];
    float d1, f1, d2, f2, ungpow, unscale;
    int32_t ntr;
    int16_t mark, shortpad, unass[14];
};
#pragma pack(pop)
static_assert(sizeof(SuHeader) == 240, "SU trace header is 240 bytes");
static_assert(offsetof(SuHeader, ns) == 114 && offsetof(SuHeader, dt) == 116, "ns / dt at bytes 114 / 116");
static_assert(offsetof(SuHeader, sx) == 72 && offsetof(SuHeader, gx) == 80, "sx / gx at bytes 72 / 80");
static_assert(offsetof(SuHeader, d1) == 180 && offsetof(SuHeader, ntr) == 204, "d1 / ntr at bytes 180 / 204");

}  // namespace

extern "C" {

int sfd_read_q_par(const char* path, sfd_q_par* p, char* err) {
    if (!path || !p) return SFD_EINVAL;
    std::memset(p, 0, sizeof *p);
    Card c;
    if (!c.open(path)) return card_fail(err, c.err);
    std::string s, e;
    char* names[9] = {p->vp_fn, p->vs_fn, p->rho_fn, p->qfp_fn, p->qfs_fn,
                      p->shot_fn1, p->shot_fn2, p->shot_fn3, p->shot_fn4};
    for (char* n : names) {
        if (!c.value_line(&s)) return card_fail(err, c.err);
        copy_name(n, s);
    }
    double v[3];
    int iv;
    // read(10,*) nshot
    if (!c.value_line(&s) || !list_reals(s, 1, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nshot = (int)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->offset_min = (float)v[0];
    p->offset_max = (float)v[1];
    if (!c.value_line(&s) || !list_reals(s, 3, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->sx0 = (float)v[0];
    p->sz0 = (float)v[1];
    p->rz0 = (float)v[2];
    // read(10,'(F)') dsx
    if (!c.value_line(&s) || !f_field(s, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dsx = (float)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nx = (int)v[0];
    p->nz = (int)v[1];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dx = (float)v[0];
    p->dz = (float)v[1];
    // read(10,'(I)') nt; read(10,'(F)') dt
    if (!c.value_line(&s) || !i_field(s, &iv, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nt = iv;
    if (!c.value_line(&s) || !f_field(s, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dt = (float)v[0];
    if (!c.value_line(&s) || !list_reals(s, 1, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->f0 = (float)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->ppxn = (float)v[0];
    p->ppxp = (float)v[1];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->ppzn = (float)v[0];
    p->ppzp = (float)v[1];
    // read(10,'(I)') fdo2, fdo1
    if (!c.value_line(&s) || !i_field(s, &iv, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->fdo2 = iv;
    if (!c.value_line(&s) || !i_field(s, &iv, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->fdo1 = iv;
    return SFD_OK;
}

int sfd_read_iso_par(const char* path, sfd_iso_par* p, char* err) {
    if (!path || !p) return SFD_EINVAL;
    std::memset(p, 0, sizeof *p);
    Card c;
    if (!c.open(path)) return card_fail(err, c.err);
    std::string s, e;
    if (!c.value_line(&s)) return card_fail(err, c.err);
    copy_name(p->mode, s);
    char* names[10] = {p->vp_fn, p->vs_fn, p->rho_fn, p->topo_fn, p->snapx_fn,
                       p->snapz_fn, p->shot_fn1, p->shot_fn2, p->shot_fn3, p->shot_fn4};
    for (char* n : names) {
        if (!c.value_line(&s)) return card_fail(err, c.err);
        copy_name(n, s);
    }
    double v[3];
    int iv;
    if (!c.value_line(&s) || !list_reals(s, 1, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nshot = (int)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->ofsmin = (float)v[0];
    p->ofsmax = (float)v[1];
    if (!c.value_line(&s) || !list_reals(s, 3, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->sx0 = (float)v[0];
    p->sz0 = (float)v[1];
    p->rz0 = (float)v[2];
    if (!c.value_line(&s) || !f_field(s, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dsx = (float)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nx = (int)v[0];
    p->nz = (int)v[1];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dx = (float)v[0];
    p->dz = (float)v[1];
    if (!c.value_line(&s) || !i_field(s, &iv, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->nt = iv;
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->dt = (float)v[0];
    p->dtsnap = (float)v[1];
    if (!c.value_line(&s) || !f_field(s, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->f0 = (float)v[0];
    if (!c.value_line(&s) || !f_field(s, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->decay = (float)v[0];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->ppxn = (float)v[0];
    p->ppxp = (float)v[1];
    if (!c.value_line(&s) || !list_reals(s, 2, v, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->ppzn = (float)v[0];
    p->ppzp = (float)v[1];
    if (!c.value_line(&s) || !i_field(s, &iv, &e)) return card_fail(err, c.err.empty() ? e : c.err);
    p->rgau = iv;
    return SFD_OK;
}

int sfd_write_su_shot(const char* path, const float* record, int ntrace, int nt, int shot_no, int src_col,
                      int first_col, float dx, float dt, int with_cdp) {
    if (!path || (!record && ntrace > 0) || ntrace < 0 || nt < 1 || nt > 65535) return SFD_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return SFD_EINVAL;
    SuHeader h;
    std::memset(&h, 0, sizeof h);
    // Real results assigned to integer fields truncate, as the Fortran
    // assignments do; the products are single precision.
    h.fldr = shot_no;
    h.trid = 1;
    h.sx = (int32_t)((float)(src_col - 1) * dx);
    h.ns = (uint16_t)nt;
    h.dt = (uint16_t)(int)(dt * 1.0e6f);
    int rc = SFD_OK;
    for (int i = 1; i <= ntrace && rc == SFD_OK; ++i) {
        h.gx = (int32_t)((float)(first_col - 1 + i - 1) * dx);
        h.offset = h.gx - h.sx;
        // Q_forward only: cdp = nint((gx + sx) / dx) + 1.
        if (with_cdp) h.cdp = (int32_t)std::lround((float)(h.gx + h.sx) / dx) + 1;
        if (std::fwrite(&h, sizeof h, 1, f) != 1 ||
            std::fwrite(record + (size_t)(i - 1) * nt, sizeof(float), nt, f) != (size_t)nt)
            rc = SFD_EINVAL;
    }
    if (std::fclose(f) != 0) rc = SFD_EINVAL;
    return rc;
}

int sfd_merge_su(const char* const* shot_files, int nshot, int ntrace, int nt, const char* out_path,
                 int remove_inputs) {
    if (!shot_files || nshot < 1 || ntrace < 1 || nt < 1 || !out_path) return SFD_EINVAL;
    const size_t bytes = (size_t)ntrace * (nt + 60) * 4;
    std::vector<char> buf(bytes);
    FILE* out = std::fopen(out_path, "wb");
    if (!out) return SFD_EINVAL;
    int rc = SFD_OK;
    for (int s = 0; s < nshot && rc == SFD_OK; ++s) {
        // Each shot file is one record of ntrace traces (the reference reads
        // it as one direct-access record of ntrace * (nt + 60) words).
        FILE* in = shot_files[s] ? std::fopen(shot_files[s], "rb") : nullptr;
        if (!in) {
            rc = SFD_EINVAL;
            break;
        }
        const size_t got = std::fread(buf.data(), 1, bytes, in);
        const bool longer = std::fgetc(in) != EOF;
        std::fclose(in);
        if (got != bytes || longer || std::fwrite(buf.data(), 1, bytes, out) != bytes) rc = SFD_EINVAL;
    }
    if (std::fclose(out) != 0) rc = SFD_EINVAL;
    if (rc == SFD_OK && remove_inputs)
        for (int s = 0; s < nshot; ++s) std::remove(shot_files[s]);
    return rc;
}

}  // extern "C"
