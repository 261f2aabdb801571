// Parser consistency check: the counting pass (count_piece) and the parsing
// pass (scan_piece) agree on every piece, and a parse with several threads
// equals the one-thread parse, over random byte strings built from the bytes
// the reader treats specially (digits, '-', '\n', CR, space, others).
#include "dataset.cpp"
#include <cstring>
#include <random>
int main() {
    std::mt19937_64 g(7);
    const char alpha[] = "0123456789-\n \ra\n--9";
    int64_t bad = 0, cases = 0;
    for (int it = 0; it < 50000; it++) {
        const int len = (int)(g() % 300);
        std::string s(len, ' ');
        for (auto &c : s) c = alpha[g() % (sizeof(alpha) - 1)];
        for (int fr = 0; fr <= 2; fr++) {
            // a piece starts at the file start or right after '\n'
            int64_t a, b, c2, d2; int32_t mx = 0;
            tfim::count_piece(s.data(), s.data() + len, fr, a, b);
            std::vector<int32_t> ids(a + 1); std::vector<int64_t> ends(b + 1);
            bool big = false;
            tfim::scan_piece(s.data(), s.data() + len, fr, ids.data(), a, ends.data(), b, 0, c2, d2, mx, big);
            cases++;
            if (a != c2 || b != d2) { if (bad++ < 5) fprintf(stderr, "mismatch fr=%d '%s' count %lld/%lld scan %lld/%lld\n", fr, s.c_str(), (long long)a, (long long)b, (long long)c2, (long long)d2); }
        }
    }
    // whole-image parse with several thread counts equals one thread
    for (int it = 0; it < 1000; it++) {
        const int len = (int)(g() % 5000);
        std::string s(len, ' ');
        for (auto &c : s) c = alpha[g() % (sizeof(alpha) - 1)];
        // digit runs of at most 8 digits here (ids within int32)
        for (int i = 0, run = 0; i < len; i++) {
            run = (s[i] >= '0' && s[i] <= '9') ? run + 1 : 0;
            if (run > 8) { s[i] = ' '; run = 0; }
        }
        tfim::IdVec i1; tfim::OffVec o1; int32_t m1 = 0;
        std::string e1, e2;
        bool ok1 = tfim::parse_parallel(s.data(), s.size(), it & 1, 1, i1, o1, m1, e1);
        for (int th = 2; th <= 9; th++) {
            tfim::IdVec i2; tfim::OffVec o2; int32_t m2 = 0;
            bool ok2 = tfim::parse_parallel(s.data(), s.size(), it & 1, th, i2, o2, m2, e2);
            if (ok1 != ok2 || (ok1 && (i1 != i2 || o1 != o2 || m1 != m2))) { if (bad++ < 10) fprintf(stderr, "parallel mismatch th=%d len=%d\n", th, len); }
        }
    }
    // ids up to kMaxItemId (INT32_MAX - 2) parse; one past it (or a longer run) is refused
    struct { const char *text; bool ok; } lim[] = {{"1 2147483645\n", true}, {"1 2147483646\n", false},
                                                  {"1 2147483647\n", false}, {"1 2147483648\n", false},
                                                  {"5 99999999999999999999999\n", false}, {"-99999999999999999999 3\n", true}};
    for (auto &t : lim) {
        tfim::IdVec i3; tfim::OffVec o3; int32_t m3 = 0; std::string e3;
        const bool ok = tfim::parse_parallel(t.text, strlen(t.text), false, 1, i3, o3, m3, e3);
        cases++;
        if (ok != t.ok) { if (bad++ < 10) fprintf(stderr, "limit case '%s': ok=%d\n", t.text, (int)ok); }
    }
    printf("cases %lld bad %lld\n", (long long)cases, (long long)bad);
    return bad != 0;
}
