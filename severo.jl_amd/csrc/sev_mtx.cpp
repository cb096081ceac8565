// MatrixMarket coordinate files (10X matrix.mtx / matrix.mtx.gz) on the host.
//
// Follows readMM (src/input.jl:26-115): the header must read
// "%%MatrixMarket matrix coordinate real|integer general" (five tokens,
// checked in the reference's order), blank lines and lines starting with '%'
// are skipped up to the size line (three integers: rows, columns, entries),
// and then exactly `entries` lines follow, each with three whitespace
// separated tokens (row, column, value; 1-based).  Lines past the entries
// are not read, as the reference stops there.  A ".gz" path is inflated with
// zlib (the reference opens it through GzipDecompressorStream); other files
// are memory-mapped.  The entries are parsed by host threads, each taking a
// range of lines (line boundaries found first, so every entry lands at its
// own index).
#include "sev_kernels.hpp"

#include <zlib.h>

#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace sev {

namespace {

// The file's bytes: a read-only mapping, or the inflated contents of a .gz.
struct MtxText {
    const char *p = nullptr;
    size_t n = 0;
    void *map = nullptr;
    size_t maplen = 0;
    std::vector<char> buf;
    MtxText() = default;
    MtxText(const MtxText &) = delete;
    MtxText &operator=(const MtxText &) = delete;
    ~MtxText() {
        if (map) munmap(map, maplen);
    }
};

bool ends_with(const std::string &s, const char *suf) {
    const size_t k = std::strlen(suf);
    return s.size() >= k && s.compare(s.size() - k, k, suf) == 0;
}

void open_text(const char *path, MtxText &t) {
    SEV_REQUIRE(path != nullptr, "path is NULL");
    const std::string ps(path);
    if (ends_with(ps, ".gz")) {
        gzFile f = gzopen(path, "rb");
        SEV_REQUIRE(f != nullptr, "cannot open " + ps);
        gzbuffer(f, 1 << 20);
        const size_t chunk = (size_t)64 << 20;
        size_t have = 0;
        for (;;) {
            t.buf.resize(have + chunk);
            const int got = gzread(f, t.buf.data() + have, (unsigned)chunk);
            if (got < 0) {
                int err = 0;
                const std::string msg = gzerror(f, &err);
                gzclose(f);
                throw std::runtime_error("cannot inflate " + ps + ": " + msg);
            }
            have += (size_t)got;
            if (got == 0) break;
        }
        gzclose(f);
        t.buf.resize(have);
        t.p = t.buf.data();
        t.n = have;
        return;
    }
    const int fd = open(path, O_RDONLY);
    SEV_REQUIRE(fd >= 0, "cannot open " + ps);
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        throw std::runtime_error("cannot stat " + ps);
    }
    t.n = (size_t)sb.st_size;
    if (t.n > 0) {
        t.map = mmap(nullptr, t.n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (t.map == MAP_FAILED) {
            t.map = nullptr;
            close(fd);
            throw std::runtime_error("cannot map " + ps);
        }
        t.maplen = t.n;
        madvise(t.map, t.n, MADV_SEQUENTIAL);
        t.p = (const char *)t.map;
    }
    close(fd);
}

// One line starting at pos (without its "\n" or "\r\n"); pos moves past it.
// At the end of the text the line is empty, as Julia's readline returns.
std::string next_line(const MtxText &t, size_t &pos) {
    if (pos >= t.n) return std::string();
    const char *b = t.p + pos;
    const void *nl = std::memchr(b, '\n', t.n - pos);
    size_t len = nl ? (size_t)((const char *)nl - b) : t.n - pos;
    pos += len + (nl ? 1 : 0);
    if (len > 0 && b[len - 1] == '\r') --len;
    return std::string(b, len);
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\v' || c == '\f' || c == '\r'; }

std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && is_ws(s[i])) ++i;
        size_t j = i;
        while (j < s.size() && !is_ws(s[j])) ++j;
        if (j > i) out.emplace_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// Int64 of [b, e): optional sign and decimal digits only (Julia's
// parse(Int64, x)).
bool parse_i64(const char *b, const char *e, int64_t &v) {
    if (b < e && *b == '+') ++b;  // from_chars takes '-' but not '+'
    if (b == e || *b == '+') return false;
    const auto r = std::from_chars(b, e, v);
    return r.ec == std::errc() && r.ptr == e;
}

// Float64 of [b, e) (decimal or "inf"/"nan" forms, locale independent).
bool parse_f64(const char *b, const char *e, double &v) {
    if (b < e && *b == '+') ++b;
    if (b == e || *b == '+') return false;
    const auto r = std::from_chars(b, e, v);
    return r.ec == std::errc() && r.ptr == e;
}

}  // namespace

struct MtxReader::Impl {
    MtxText text;
    size_t data_pos = 0;  // first byte after the size line
};

MtxReader::MtxReader(const char *path) : impl_(new Impl) {
    MtxText &t = impl_->text;
    open_text(path, t);
    size_t pos = 0;
    const std::string header = next_line(t, pos);
    const std::vector<std::string> tok = split_ws(header);
    if (tok.size() != 5) throw std::runtime_error("Malformed MatrixMarket header: " + header);
    if (tok[0] != "%%MatrixMarket")
        throw std::runtime_error("Not a valid MatrixMarket header: " + header);
    const std::string obj = lower(tok[1]), format = lower(tok[2]), field = lower(tok[3]),
                      symm = lower(tok[4]);
    if (obj != "matrix")
        throw std::runtime_error("Unknown MatrixMarket data type: " + obj +
                                 " (only \"matrix\" is supported)");
    if (field == "real")
        real_ = true;
    else if (field == "integer")
        real_ = false;
    else
        throw std::runtime_error("Unsupported field " + field +
                                 " (only real and integer are supported)");
    if (symm != "general")
        throw std::runtime_error("Unknown matrix symmetry: " + symm +
                                 " (only \"general\" is supported)");
    if (format != "coordinate")
        throw std::runtime_error("Unsupported matrix format: " + format +
                                 " (only \"coordinate\" is supported)");
    std::string ll = next_line(t, pos);
    while (ll.empty() || ll[0] == '%') {
        if (pos >= t.n && ll.empty()) break;
        ll = next_line(t, pos);
    }
    const std::vector<std::string> dd = split_ws(ll);
    int64_t d[3] = {0, 0, 0};
    bool ok = dd.size() == 3;
    for (size_t i = 0; ok && i < 3; ++i)
        ok = parse_i64(dd[i].data(), dd[i].data() + dd[i].size(), d[i]);
    if (!ok) throw std::runtime_error("Could not read in matrix dimensions: " + ll);
    SEV_REQUIRE(d[0] >= 0 && d[1] >= 0 && d[2] >= 0, "negative matrix dimensions");
    rows_ = d[0];
    cols_ = d[1];
    entries_ = d[2];
    impl_->data_pos = pos;
}

MtxReader::~MtxReader() = default;

void MtxReader::parse(const std::function<void(int64_t, int64_t, int64_t, int64_t, double)> &emit,
                      const std::function<void(int64_t first, int64_t count, int thread)> &done) {
    const MtxText &t = impl_->text;
    const size_t beg = impl_->data_pos, end = t.n;
    const int64_t E = entries_;
    if (E == 0) return;
    const int64_t nt = std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)std::thread::hardware_concurrency(), 16,
                              (int64_t)((end - beg) >> 22) + 1}));
    // chunk starts at line boundaries
    std::vector<size_t> cb((size_t)nt + 1, end);
    cb[0] = beg;
    for (int64_t c = 1; c < nt; ++c) {
        size_t s = beg + (end - beg) * (size_t)c / (size_t)nt;
        if (s < cb[(size_t)c - 1]) s = cb[(size_t)c - 1];
        const void *nl = s < end ? std::memchr(t.p + s, '\n', end - s) : nullptr;
        cb[(size_t)c] = nl ? (size_t)((const char *)nl - t.p) + 1 : end;
    }
    // lines per chunk (a last line without "\n" counts)
    std::vector<int64_t> nl((size_t)nt, 0);
    auto count_lines = [&](int64_t c) {
        const char *p = t.p + cb[(size_t)c], *e = t.p + cb[(size_t)c + 1];
        int64_t k = 0;
        while (p < e) {
            const void *q = std::memchr(p, '\n', (size_t)(e - p));
            ++k;
            if (!q) break;
            p = (const char *)q + 1;
        }
        nl[(size_t)c] = k;
    };
    {
        std::vector<std::thread> th;
        for (int64_t c = 1; c < nt; ++c) th.emplace_back(count_lines, c);
        count_lines(0);
        for (auto &x : th) x.join();
    }
    std::vector<int64_t> first((size_t)nt + 1, 0);
    for (int64_t c = 0; c < nt; ++c) first[(size_t)c + 1] = first[(size_t)c] + nl[(size_t)c];
    if (first[(size_t)nt] < E) throw std::runtime_error("Could not read matrix entry: ");
    std::vector<std::string> err((size_t)nt);
    std::vector<int64_t> err_line((size_t)nt, -1);
    auto work = [&](int64_t c) {
        const int64_t i0 = first[(size_t)c];
        if (i0 >= E) return;
        const char *p = t.p + cb[(size_t)c], *e = t.p + cb[(size_t)c + 1];
        int64_t i = i0;
        while (p < e && i < E) {
            const char *q = (const char *)std::memchr(p, '\n', (size_t)(e - p));
            const char *le = q ? q : e;
            const char *ls = p;
            p = q ? q + 1 : e;
            // three tokens
            const char *tb[3], *te[3];
            int nt3 = 0;
            const char *s = ls;
            bool extra = false;
            while (s < le) {
                while (s < le && is_ws(*s)) ++s;
                if (s >= le) break;
                const char *w = s;
                while (s < le && !is_ws(*s)) ++s;
                if (nt3 < 3) {
                    tb[nt3] = w;
                    te[nt3] = s;
                }
                ++nt3;
                if (nt3 > 3) {
                    extra = true;
                    break;
                }
            }
            int64_t r = 0, cc = 0, vi = 0;
            double vf = 0.0;
            bool ok = nt3 == 3 && !extra;
            if (!ok) {
                err[(size_t)c] = "Could not read matrix entry: ";
            } else {
                ok = parse_i64(tb[0], te[0], r) && parse_i64(tb[1], te[1], cc) &&
                     (real_ ? parse_f64(tb[2], te[2], vf) : parse_i64(tb[2], te[2], vi));
                if (!ok) err[(size_t)c] = "Could not parse matrix entry: ";
            }
            if (!ok) {
                size_t len = (size_t)(le - ls);
                if (len > 0 && ls[len - 1] == '\r') --len;
                err[(size_t)c] += std::string(ls, std::min<size_t>(len, 200));
                err_line[(size_t)c] = i;
                return;
            }
            emit(i, r, cc, vi, vf);
            ++i;
        }
        if (done) done(i0, i - i0, (int)c);
    };
    {
        std::vector<std::thread> th;
        for (int64_t c = 1; c < nt; ++c) th.emplace_back(work, c);
        work(0);
        for (auto &x : th) x.join();
    }
    // the first failing line in file order
    int64_t best = -1;
    for (int64_t c = 0; c < nt; ++c)
        if (err_line[(size_t)c] >= 0 && (best < 0 || err_line[(size_t)c] < err_line[(size_t)best]))
            best = c;
    if (best >= 0) throw std::runtime_error(err[(size_t)best]);
}

}  // namespace sev

// ---------------------------------------------------------------------------
// C ABI: the host reader (readMM's (I, J, V) vectors, before sparse()).
// ---------------------------------------------------------------------------
extern "C" {

int sev_mtx_info(const char *path, int64_t *rows, int64_t *cols, int64_t *entries,
                 int32_t *is_real) {
    return sev::guarded_fn([&] {
        SEV_REQUIRE(rows && cols && entries && is_real, "output pointer is NULL");
        sev::MtxReader rd(path);
        *rows = rd.rows();
        *cols = rd.cols();
        *entries = rd.entries();
        *is_real = rd.is_real() ? 1 : 0;
    });
}

int sev_mtx_read(const char *path, int64_t *I, int64_t *J, int64_t *Vi, double *Vf) {
    return sev::guarded_fn([&] {
        sev::MtxReader rd(path);
        const int64_t E = rd.entries();
        SEV_REQUIRE(E == 0 || (I && J), "I or J is NULL");
        SEV_REQUIRE(E == 0 || (rd.is_real() ? Vf != nullptr : Vi != nullptr),
                    rd.is_real() ? "Vf is NULL (real field)" : "Vi is NULL (integer field)");
        const bool real = rd.is_real();
        rd.parse(
            [&](int64_t i, int64_t r, int64_t c, int64_t vi, double vf) {
                I[i] = r;
                J[i] = c;
                if (real)
                    Vf[i] = vf;
                else
                    Vi[i] = vi;
            },
            nullptr);
    });
}

}  // extern "C"
