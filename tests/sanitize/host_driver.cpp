// The library's host-only parsers and writers built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile): the parameter-card
// readers and SU writers of io.cpp and the tuning-cache JSON of tuning.cpp,
// driven over the reference's own cards, every truncation of them, SU files
// written and merged, and the JSON cache saved, reloaded and corrupted byte
// by byte.  Exit status 0 = every call returned without a sanitizer report
// and with the expected status.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"

// The library's error helper (context.cpp, which also holds device code and
// is not part of this host build).
namespace sfd {
int fail(sfd_context* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}
}  // namespace sfd

static int failures = 0;
#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                              \
        }                                                            \
    } while (0)

static std::string slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

static void spit(const std::string& p, const std::string& s) {
    std::ofstream f(p, std::ios::binary);
    f << s;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string gold = argv[1], tmp = argv[2];
    char err[256];
    // Parameter cards: the reference's own, then every truncation (the reader
    // must return SFD_OK or SFD_EINVAL, nothing else, and never overrun).
    sfd_q_par q;
    CHECK(sfd_read_q_par((gold + "/isoq_2d_modelling_simple.par").c_str(), &q, err) == SFD_OK);
    sfd_iso_par iso;
    CHECK(sfd_read_iso_par((gold + "/iso_2d_modelling_sin_test.par").c_str(), &iso, err) == SFD_OK);
    for (const char* card : {"isoq_2d_modelling_simple.par", "iso_2d_modelling_sin_test.par"}) {
        const std::string text = slurp(gold + "/" + card);
        const bool isq = card[3] == 'q';
        for (size_t cut = 0; cut < text.size(); cut += 3) {
            spit(tmp + "/cut.par", text.substr(0, cut));
            const int rc = isq ? sfd_read_q_par((tmp + "/cut.par").c_str(), &q, err)
                               : sfd_read_iso_par((tmp + "/cut.par").c_str(), &iso, err);
            CHECK(rc == SFD_OK || rc == SFD_EINVAL);
        }
        // one very long line in place of every value line
        std::string junk(5000, '7');
        spit(tmp + "/long.par", text + "\n" + junk + "\n");
        const int rc = isq ? sfd_read_q_par((tmp + "/long.par").c_str(), &q, err)
                           : sfd_read_iso_par((tmp + "/long.par").c_str(), &iso, err);
        CHECK(rc == SFD_OK || rc == SFD_EINVAL);
    }
    CHECK(sfd_read_q_par((tmp + "/missing.par").c_str(), &q, err) == SFD_EINVAL);
    // SU shot files and their merge.
    const int ntrace = 7, nt = 33;
    std::vector<float> rec((size_t)ntrace * nt);
    for (size_t i = 0; i < rec.size(); ++i) rec[i] = (float)i;
    std::vector<std::string> names;
    for (int s = 0; s < 3; ++s) {
        names.push_back(tmp + "/shot" + std::to_string(s) + ".su");
        CHECK(sfd_write_su_shot(names.back().c_str(), rec.data(), ntrace, nt, s + 1, 10 + s, 3, 12.5f, 0.001f, s & 1) ==
              SFD_OK);
    }
    std::vector<const char*> ptrs;
    for (auto& n : names) ptrs.push_back(n.c_str());
    CHECK(sfd_merge_su(ptrs.data(), 3, ntrace, nt, (tmp + "/all.su").c_str(), 0) == SFD_OK);
    CHECK(slurp(tmp + "/all.su").size() == 3 * (size_t)ntrace * (240 + 4 * nt));
    CHECK(sfd_merge_su(ptrs.data(), 3, ntrace + 1, nt, (tmp + "/bad.su").c_str(), 0) != SFD_OK);
    // Tuning cache: save, reload, then every truncation and byte corruption.
    sfd_context a;
    a.arch = "gfx950:sramecc+:xnack-";
    a.tuning["acoustic:R8:cpml:x12:z11"] = sfd::TunedShape{0, 16, 4, 57.2f};
    a.tuning["batch:R8:cpml:n6:c27"] = sfd::TunedShape{2, 32, 8, 1000.f};
    const std::string jp = tmp + "/tune.json";
    CHECK(sfd_context_save_tuning(&a, jp.c_str()) == SFD_OK);
    sfd_context b;
    b.arch = a.arch;
    CHECK(sfd_context_load_tuning(&b, jp.c_str()) == SFD_OK);
    CHECK(sfd_context_tuning_size(&b) == 2);
    const std::string js = slurp(jp);
    for (size_t cut = 0; cut + 1 < js.size(); ++cut) {
        spit(tmp + "/t.json", js.substr(0, cut));
        sfd_context c;
        c.arch = a.arch;
        const int rc = sfd_context_load_tuning(&c, (tmp + "/t.json").c_str());
        CHECK(rc == SFD_OK || rc == SFD_EINVAL);
    }
    for (size_t i = 0; i < js.size(); ++i) {
        for (char ch : {'"', '{', '[', '0', '\\', '\0', '-'}) {
            std::string bad = js;
            bad[i] = ch;
            spit(tmp + "/t.json", bad);
            sfd_context c;
            c.arch = a.arch;
            const int rc = sfd_context_load_tuning(&c, (tmp + "/t.json").c_str());
            CHECK(rc == SFD_OK || rc == SFD_EINVAL);
        }
    }
    printf("host driver: %d failures\n", failures);
    return failures ? 1 : 0;
}
