// tfim_verify -- compares Results files the way the reference's checker does
// (scripts/verify_all.py:13-23,40): each line matching
//   <n>th: <len> { <ids> }  support: <s> frequency: <f>
// contributes {set of ids -> f}, later lines overwriting earlier ones with the
// same set; two files agree iff these maps are equal.  --header also compares
// the header fields that pin the reduction (transNum, dimensionReduced,
// transNumReduced, MFIsNumber).  Needs no Python on the machine that runs it.
//
//   tfim_verify <results> <gold> [--header]
//   tfim_verify --dir <results_dir> <gold_dir> [--header]
//       every "*=Results.txt" in results_dir against the same name in gold_dir
//
// Prints "[identical] ..." / "[MISMATCH] ..." per file (with up to 5 sets that
// differ) and exits 0 iff every compared file agrees (1 otherwise, 2 on usage
// or unreadable input).
#include <dirent.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

using MfiMap = std::map<std::vector<long long>, long long>;

bool skip_ws(const char *&p, bool need) {
    const char *s = p;
    while (*p && isspace((unsigned char)*p)) p++;
    return !need || p > s;
}
bool digits(const char *&p, long long *v) {
    const char *s = p;
    long long x = 0;
    while (*p >= '0' && *p <= '9') x = x * 10 + (*p++ - '0');
    if (v) *v = x;
    return p > s;
}
bool lit(const char *&p, const char *w) {
    const size_t n = strlen(w);
    if (strncmp(p, w, n)) return false;
    p += n;
    return true;
}

// One line; true when it is an MFI line.
bool parse_line(const char *ln, std::vector<long long> &items, long long &freq) {
    const char *p = ln;
    skip_ws(p, false);
    if (!digits(p, nullptr) || !lit(p, "th:") || !skip_ws(p, true) || !digits(p, nullptr) || !skip_ws(p, true) ||
        !lit(p, "{"))
        return false;
    const char *close = strchr(p, '}');
    if (!close) return false;
    items.clear();
    std::string body(p, close);
    const char *b = body.c_str();
    char *end = nullptr;
    while (true) {
        while (*b && isspace((unsigned char)*b)) b++;
        if (!*b) break;
        const long long v = strtoll(b, &end, 10);
        if (end == b) return false;   // the checker's int() would raise
        items.push_back(v);
        b = end;
        if (*b && !isspace((unsigned char)*b)) return false;
    }
    p = close + 1;
    if (!skip_ws(p, true) || !lit(p, "support:") || !skip_ws(p, true)) return false;
    const char *s = p;
    while (*p && !isspace((unsigned char)*p)) p++;
    if (p == s || !skip_ws(p, true) || !lit(p, "frequency:") || !skip_ws(p, true) || !digits(p, &freq)) return false;
    std::sort(items.begin(), items.end());
    items.erase(std::unique(items.begin(), items.end()), items.end());
    return true;
}

bool load(const std::string &path, MfiMap &m, std::map<std::string, std::string> &hdr) {
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    std::string line;
    char buf[1 << 16];
    bool in_header = true;
    std::vector<long long> items;
    long long freq = 0;
    while (fgets(buf, sizeof buf, f)) {
        line.assign(buf);
        while (!line.empty() && line.back() != '\n' && fgets(buf, sizeof buf, f)) line += buf;
        if (in_header && line.find("List of MFIs") != std::string::npos) in_header = false;
        if (in_header) {
            const size_t c = line.find(':');
            if (c != std::string::npos) {
                auto trim = [](std::string x) {
                    while (!x.empty() && isspace((unsigned char)x.back())) x.pop_back();
                    size_t i = 0;
                    while (i < x.size() && isspace((unsigned char)x[i])) i++;
                    return x.substr(i);
                };
                const std::string k = trim(line.substr(0, c));
                if (!hdr.count(k)) hdr[k] = trim(line.substr(c + 1));
            }
        }
        if (parse_line(line.c_str(), items, freq)) m[items] = freq;
    }
    fclose(f);
    return true;
}

std::string show(const std::vector<long long> &s) {
    std::string r = "{";
    for (size_t i = 0; i < s.size(); i++) r += (i ? " " : "") + std::to_string(s[i]);
    return r + "}";
}

// 0 identical, 1 mismatch, 2 unreadable
int compare(const std::string &mine, const std::string &gold, const std::string &name, bool header) {
    MfiMap a, b;
    std::map<std::string, std::string> ha, hb;
    if (!load(mine, a, ha) || !load(gold, b, hb)) {
        printf("[unreadable] %s\n", name.c_str());
        return 2;
    }
    std::vector<std::string> bad;
    int shown = 0;
    for (const auto &kv : a) {
        auto it = b.find(kv.first);
        if (it == b.end() || it->second != kv.second) {
            if (shown++ < 5)
                bad.push_back("  mine " + show(kv.first) + " frequency " + std::to_string(kv.second) +
                              (it == b.end() ? " (not in gold)" : " (gold " + std::to_string(it->second) + ")"));
        }
    }
    for (const auto &kv : b)
        if (!a.count(kv.first) && shown++ < 5)
            bad.push_back("  gold " + show(kv.first) + " frequency " + std::to_string(kv.second) + " (missing)");
    if (header)
        for (const char *k : {"transNum", "dimensionReduced", "transNumReduced", "MFIsNumber"})
            if (ha[k] != hb[k]) bad.push_back(std::string("  header ") + k + ": " + ha[k] + " vs " + hb[k]);
    if (bad.empty() && a == b) {
        printf("[identical] %s (%zu MFIs)\n", name.c_str(), b.size());
        return 0;
    }
    printf("[MISMATCH] %s (mine=%zu gold=%zu)\n", name.c_str(), a.size(), b.size());
    for (const auto &s : bad) printf("%s\n", s.c_str());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    bool header = false, dir = false;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--header")) header = true;
        else if (!strcmp(argv[i], "--dir")) dir = true;
        else pos.push_back(argv[i]);
    }
    if (pos.size() != 2) {
        fprintf(stderr, "usage: tfim_verify <results> <gold> [--header] | --dir <results_dir> <gold_dir> [--header]\n");
        return 2;
    }
    if (!dir) {
        const int r = compare(pos[0], pos[1], pos[0], header);
        return r;
    }
    DIR *d = opendir(pos[0].c_str());
    if (!d) {
        fprintf(stderr, "cannot read %s\n", pos[0].c_str());
        return 2;
    }
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) {
        const std::string n = e->d_name;
        const std::string suf = "=Results.txt";
        if (n.size() > suf.size() && n.compare(n.size() - suf.size(), suf.size(), suf) == 0) names.push_back(n);
    }
    closedir(d);
    std::sort(names.begin(), names.end());
    int ok = 0, bad = 0;
    for (const auto &n : names) {
        const std::string g = pos[1] + "/" + n;
        FILE *f = fopen(g.c_str(), "r");
        if (!f) {
            printf("[no gold] %s\n", n.c_str());
            continue;
        }
        fclose(f);
        const int r = compare(pos[0] + "/" + n, g, n, header);
        (r == 0 ? ok : bad)++;
    }
    printf("\n%d identical, %d mismatched\n", ok, bad);
    return bad ? 1 : 0;
}
