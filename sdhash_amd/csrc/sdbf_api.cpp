// C++ classes sdbf / sdbf_set with the reference's interface
// (sdbf/sdbf_class.h, sdbf/sdbf_set.h) on top of the C ABI of the gfx950
// engine.  Digesting and comparing run on the device; this file holds the
// host-side parts: object life cycle, the text form (to_string and its
// parser, sdbf/sdbf_class.cc:144-215,294-334) and result formatting
// (sdbf/sdbf_set.cc:240-394).
#include "sdbf_amd.h"

#include <sys/stat.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace {

const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

// RFC 4648 base64 with '=' padding (the encoding of base64/modp_b64.cc)
std::string b64encode(const uint8_t *in, size_t len)
{
    std::string out;
    out.reserve((len + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < len; i += 3) {
        const uint32_t v = (uint32_t)in[i] << 16 | (uint32_t)in[i + 1] << 8 | in[i + 2];
        out += kB64[v >> 18 & 63];
        out += kB64[v >> 12 & 63];
        out += kB64[v >> 6 & 63];
        out += kB64[v & 63];
    }
    if (len - i == 1) {
        const uint32_t v = (uint32_t)in[i] << 16;
        out += kB64[v >> 18 & 63];
        out += kB64[v >> 12 & 63];
        out += "==";
    } else if (len - i == 2) {
        const uint32_t v = (uint32_t)in[i] << 16 | (uint32_t)in[i + 1] << 8;
        out += kB64[v >> 18 & 63];
        out += kB64[v >> 12 & 63];
        out += kB64[v >> 6 & 63];
        out += '=';
    }
    return out;
}

int b64val(char c)
{
    if (c >= 'A' && c <= 'Z')
        return c - 'A';
    if (c >= 'a' && c <= 'z')
        return c - 'a' + 26;
    if (c >= '0' && c <= '9')
        return c - '0' + 52;
    if (c == '+')
        return 62;
    if (c == '/')
        return 63;
    return -1;
}

// decodes len characters (a multiple of 4, '=' padding at the end only);
// returns the decoded length or -1 on malformed input
long b64decode(const char *in, size_t len, std::vector<uint8_t> &out)
{
    out.clear();
    if (len % 4)
        return -1;
    for (size_t i = 0; i < len; i += 4) {
        int v[4];
        int pad = 0;
        for (int k = 0; k < 4; k++) {
            if (in[i + k] == '=' && i + 4 == len && k >= 2) {
                v[k] = 0;
                pad++;
            } else {
                if (pad)
                    return -1;
                v[k] = b64val(in[i + k]);
                if (v[k] < 0)
                    return -1;
            }
        }
        const uint32_t x = (uint32_t)v[0] << 18 | (uint32_t)v[1] << 12 | (uint32_t)v[2] << 6 | (uint32_t)v[3];
        out.push_back((uint8_t)(x >> 16));
        if (pad < 2)
            out.push_back((uint8_t)(x >> 8));
        if (pad < 1)
            out.push_back((uint8_t)x);
    }
    return (long)out.size();
}

char *dupname(const char *s)
{
    if (!s)
        s = "";
    const size_t n = strlen(s) + 1;
    char *d = (char *)malloc(n);
    memcpy(d, s, n);
    return d;
}

void check(sdbfx_ctx *ctx, int rc)
{
    if (rc != 0) {
        if (sdbf::config->warnings)
            std::cerr << "sdbf engine: " << sdbfx_last_error(ctx) << std::endl;
        throw -1;
    }
}

}  // namespace

// ----------------------------------------------------------------- sdbf_conf

sdbf_conf::sdbf_conf(uint32_t thread_cnt_, uint32_t warnings_, uint32_t max_elem_ct, uint32_t max_elem_ct_dd)
    : thread_cnt(thread_cnt_), max_elem(max_elem_ct), max_elem_dd(max_elem_ct_dd), warnings(warnings_)
{
    if (const char *d = getenv("SDBFX_DEVICE"))
        device = atoi(d);
}

sdbf_conf::~sdbf_conf()
{
    if (ctx_)
        sdbfx_ctx_destroy(ctx_);
}

sdbfx_ctx *sdbf_conf::ctx()
{
    if (!ctx_ && sdbfx_ctx_create(device, 0, &ctx_) != 0) {
        std::cerr << "sdbf engine: cannot open HIP device " << device << std::endl;
        throw -1;
    }
    return ctx_;
}

sdbf_conf *sdbf::config = new sdbf_conf(1, 0, 160, 192);

// ----------------------------------------------------------------- sdbf

void sdbf::from_device(sdbfx_set *set, uint32_t idx, const char *name)
{
    sdbfx_digest_info info;
    if (sdbfx_set_info(set, idx, &info) != 0)
        throw -1;
    hashname = dupname(name ? name : info.name);
    bf_count = info.bf_count;
    last_count = info.last_count;
    dd_block_size = info.dd_block_size;
    max_elem = info.max_elem;
    orig_file_size = info.orig_size;
    buffer = (uint8_t *)malloc((size_t)bf_count * bf_size);
    hamming = (uint16_t *)malloc(bf_count * sizeof(uint16_t));
    std::vector<uint16_t> ec(bf_count);
    check(config->ctx(), sdbfx_set_read(config->ctx(), set, idx, buffer, ec.data(), hamming));
    if (dd_block_size) {
        elem_counts = (uint16_t *)malloc(bf_count * sizeof(uint16_t));
        memcpy(elem_counts, ec.data(), bf_count * sizeof(uint16_t));
    }
}

void sdbf::hash_buffer(const char *name, const uint8_t *buf, uint64_t len, uint32_t dd)
{
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    sdbfx_set *set = nullptr;
    check(ctx, sdbfx_set_create(ctx, &set));
    int32_t status = 0;
    const char *nm = name ? name : "";
    const int rc = sdbfx_set_digest(ctx, set, 1, &buf, &len, &nm, &dd, &status);
    if (rc != 0 || status != 0) {
        sdbfx_set_destroy(set);
        if (rc != 0)
            check(ctx, rc);
        throw (int)status;
    }
    from_device(set, 0, nm);
    sdbfx_set_destroy(set);
}

void sdbf::hash_batch(const std::vector<std::string> &names, const std::vector<const uint8_t *> &bufs,
                      const std::vector<uint64_t> &lens, uint32_t dd, std::vector<sdbf *> &out)
{
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    const uint32_t n = (uint32_t)bufs.size();
    std::vector<const char *> nm(n);
    std::vector<uint32_t> ddv(n, dd);
    std::vector<int32_t> status(n);
    for (uint32_t i = 0; i < n; i++)
        nm[i] = names[i].c_str();
    sdbfx_set *set = nullptr;
    check(ctx, sdbfx_set_create(ctx, &set));
    const int rc = sdbfx_set_digest(ctx, set, n, bufs.data(), lens.data(), nm.data(), ddv.data(), status.data());
    if (rc != 0) {
        sdbfx_set_destroy(set);
        check(ctx, rc);
    }
    out.assign(n, nullptr);
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (status[i] != 0)
            continue;
        sdbf *s = new sdbf();
        s->from_device(set, k++, nm[i]);
        out[i] = s;
    }
    sdbfx_set_destroy(set);
}

// sdbf/sdbf_class.cc:34-58: missing/unreadable/too-small files throw -1
sdbf::sdbf(const char *filename, uint32_t dd)
{
    struct stat st;
    if (!filename || stat(filename, &st) != 0 || !S_ISREG(st.st_mode) || st.st_size < 512)
        throw -1;
    std::ifstream is(filename, std::ios::binary);
    std::vector<uint8_t> data((size_t)st.st_size);
    is.read((char *)data.data(), (std::streamsize)data.size());
    if ((uint64_t)is.gcount() != (uint64_t)st.st_size)
        throw -1;
    hash_buffer(filename, data.data(), data.size(), dd);
}

// sdbf/sdbf_class.cc:71-102: reads msize bytes; fewer than 512 throw -3; the
// digest covers all msize bytes (zero filled past the read)
sdbf::sdbf(const char *name, std::istream *ifs, uint32_t dd, uint64_t msize, index_info *info)
{
    if (info)
        throw -1;
    std::vector<uint8_t> data(msize, 0);
    ifs->read((char *)data.data(), (std::streamsize)msize);
    const uint64_t got = (uint64_t)ifs->gcount();
    if (got < 512)
        throw -3;
    hash_buffer(name, data.data(), msize, dd);
    orig_file_size = got;
}

// sdbf/sdbf_class.cc:115-136
sdbf::sdbf(const char *name, char *str, uint32_t dd, uint64_t length, index_info *info)
{
    if (info)
        throw -1;
    if (length < 512)
        throw -3;
    hash_buffer(name, (const uint8_t *)str, length, dd);
}

// sdbf/sdbf_class.cc:144-215
sdbf::sdbf(const std::string &str)
{
    if (str.length() < 344)
        throw -1;
    std::stringstream ss(str);
    std::string magic, f;
    std::getline(ss, magic, ':');
    if (magic.compare(0, 4, "sdbf"))
        throw -2;
    try {
        std::getline(ss, f, ':');   // version
        std::getline(ss, f, ':');   // name length
        std::getline(ss, f, ':');
        hashname = dupname(f.c_str());
        std::getline(ss, f, ':');
        orig_file_size = std::stoull(f);
        std::getline(ss, f, ':');   // hash
        std::getline(ss, f, ':');
        bf_size = (uint32_t)std::stoul(f);
        std::getline(ss, f, ':');
        hash_count = (uint32_t)std::stoul(f);
        std::getline(ss, f, ':');   // mask (always 7ff)
        std::getline(ss, f, ':');
        max_elem = (uint32_t)std::stoul(f);
        std::getline(ss, f, ':');
        bf_count = (uint32_t)std::stoul(f);
    } catch (const std::exception &) {
        throw -2;
    }
    if (bf_size != 256 || bf_count == 0)
        throw -2;
    buffer = (uint8_t *)calloc(bf_count, bf_size);
    std::vector<uint8_t> dec;
    if (magic == "sdbf-dd") {
        try {
            std::getline(ss, f, ':');
            dd_block_size = (uint32_t)std::stoul(f);
        } catch (const std::exception &) {
            throw -2;
        }
        elem_counts = (uint16_t *)calloc(bf_count, sizeof(uint16_t));
        for (uint32_t i = 0; i < bf_count; i++) {
            std::getline(ss, f, ':');
            unsigned long ec = strtoul(f.c_str(), nullptr, 16);
            elem_counts[i] = (uint16_t)ec;
            std::getline(ss, f, ':');
            if (f.size() < 344 || b64decode(f.data(), 344, dec) != 256)
                throw -2;
            memcpy(buffer + (size_t)i * bf_size, dec.data(), 256);
        }
    } else {
        try {
            std::getline(ss, f, ':');
            last_count = (uint32_t)std::stoul(f);
        } catch (const std::exception &) {
            throw -2;
        }
        const uint64_t raw = (uint64_t)bf_count * bf_size;
        const uint64_t b64_len = 4 * (raw / 3 + (raw % 3 ? 1 : 0));
        std::string b64(b64_len, '\0');
        ss.read(&b64[0], (std::streamsize)b64_len);
        if ((uint64_t)ss.gcount() != b64_len || b64decode(b64.data(), b64_len, dec) != (long)raw)
            throw -2;
        memcpy(buffer, dec.data(), raw);
    }
    hamming = (uint16_t *)malloc(bf_count * sizeof(uint16_t));
    for (uint32_t i = 0; i < bf_count; i++) {
        uint32_t h = 0;
        for (uint32_t b = 0; b < bf_size; b++)
            h += (uint32_t)__builtin_popcount(buffer[(size_t)i * bf_size + b]);
        hamming[i] = (uint16_t)h;
    }
}

sdbf::~sdbf()
{
    free(buffer);
    free(hamming);
    free(elem_counts);
    free(hashname);
}

const char *sdbf::name() { return hashname; }
uint64_t sdbf::size() { return (uint64_t)bf_size * bf_count; }
uint64_t sdbf::input_size() { return orig_file_size; }
uint32_t sdbf::filter_count() { return bf_count; }

uint8_t *sdbf::clone_filter(uint32_t position)
{
    if (position >= bf_count)
        return nullptr;
    uint8_t *f = (uint8_t *)malloc(bf_size);
    memcpy(f, buffer + (size_t)position * bf_size, bf_size);
    return f;
}

int32_t sdbf::get_elem_count(sdbf *mine, uint64_t index)
{
    if (!mine->elem_counts)
        return (index < (uint64_t)mine->bf_count - 1) ? (int32_t)mine->max_elem : (int32_t)mine->last_count;
    return mine->elem_counts[index];
}

// sdbf/sdbf_class.cc:266-271 (sample is not supported: all filters are used)
int32_t sdbf::compare(sdbf *other, uint32_t sample)
{
    (void)sample;
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    sdbfx_set *set = nullptr;
    check(ctx, sdbfx_set_create(ctx, &set));
    sdbf *two[2] = {this, other};
    for (sdbf *s : two) {
        int rc;
        if (s->elem_counts)
            rc = sdbfx_set_add(ctx, set, s->hashname, s->orig_file_size, s->bf_count, 0, s->dd_block_size,
                               s->max_elem, s->buffer, s->elem_counts);
        else
            rc = sdbfx_set_add(ctx, set, s->hashname, s->orig_file_size, s->bf_count, s->last_count, 0,
                               s->max_elem, s->buffer, nullptr);
        if (rc != 0) {
            sdbfx_set_destroy(set);
            check(ctx, rc);
        }
    }
    int32_t score = -1;
    const int rc = sdbfx_compare_all(ctx, set, &score);
    sdbfx_set_destroy(set);
    check(ctx, rc);
    return score;
}

// sdbf/sdbf_class.cc:294-334
std::string sdbf::to_string() const
{
    std::stringstream hash;
    hash.fill('0');
    if (!elem_counts) {
        hash << "sdbf:" << std::setw(2) << 3 << ":";
        hash << (int)strlen(hashname) << ":" << hashname << ":" << orig_file_size << ":sha1:";
        hash << bf_size << ":" << hash_count << ":" << std::hex << mask << ":" << std::dec;
        hash << max_elem << ":" << bf_count << ":" << last_count << ":";
        hash << b64encode(buffer, (size_t)bf_count * bf_size);
    } else {
        hash << "sdbf-dd:" << std::setw(2) << 3 << ":";
        hash << (int)strlen(hashname) << ":" << hashname << ":" << orig_file_size << ":sha1:";
        hash << bf_size << ":" << hash_count << ":" << std::hex << mask << ":" << std::dec;
        hash << max_elem << ":" << bf_count << ":" << dd_block_size;
        for (uint32_t i = 0; i < bf_count; i++) {
            hash << ":" << std::setw(2) << std::hex << elem_counts[i];
            hash << ":" << b64encode(buffer + (size_t)i * bf_size, bf_size);
        }
    }
    hash << std::endl;
    return hash.str();
}

std::ostream &operator<<(std::ostream &os, const sdbf &s) { return os << s.to_string(); }
std::ostream &operator<<(std::ostream &os, const sdbf *s) { return os << s->to_string(); }

// ----------------------------------------------------------------- sdbf_set

sdbf_set::sdbf_set() {}

// sdbf/sdbf_set.cc:51-74
sdbf_set::sdbf_set(const char *fname)
{
    struct stat st;
    if (fname && stat(fname, &st) == 0 && S_ISREG(st.st_mode)) {
        std::ifstream ifs(fname, std::ios::binary);
        std::string line;
        while (std::getline(ifs, line)) {
            if (line.empty())
                break;
            setname = fname;
            items.push_back(new sdbf(line));
        }
    }
}

sdbf_set::~sdbf_set()
{
    if (dev)
        sdbfx_set_destroy(dev);
}

sdbf *sdbf_set::at(uint32_t pos) { return pos < items.size() ? items[pos] : nullptr; }

void sdbf_set::add(sdbf *hash)
{
    std::lock_guard<std::mutex> g(add_hash_mutex);
    items.push_back(hash);
}

void sdbf_set::add(sdbf_set *hashset)
{
    for (sdbf *s : hashset->items)
        items.push_back(s);
}

uint64_t sdbf_set::size() { return items.size(); }

uint64_t sdbf_set::input_size()
{
    uint64_t n = 0;
    for (sdbf *s : items)
        n += s->input_size();
    return n;
}

uint64_t sdbf_set::filter_count()
{
    uint64_t n = 0;
    for (sdbf *s : items)
        n += s->filter_count();
    return n;
}

int sdbf_set::empty() { return items.empty() ? 1 : 0; }

std::string sdbf_set::to_string() const
{
    std::stringstream b;
    for (const sdbf *s : items)
        b << s;
    return b.str();
}

std::ostream &operator<<(std::ostream &os, const sdbf_set &s) { return os << s.to_string(); }
std::ostream &operator<<(std::ostream &os, const sdbf_set *s) { return os << s->to_string(); }

std::string sdbf_set::name() const { return setname; }
void sdbf_set::set_name(std::string name) { setname = name; }
void sdbf_set::set_separator(char s) { sep = s; }

sdbfx_set *sdbf_set::device_set()
{
    sdbfx_ctx *ctx = sdbf::config->ctx();
    if (dev && dev_items > items.size()) {
        sdbfx_set_destroy(dev);
        dev = nullptr;
    }
    if (!dev) {
        check(ctx, sdbfx_set_create(ctx, &dev));
        dev_items = 0;
    }
    for (; dev_items < items.size(); dev_items++) {
        sdbf *s = items[dev_items];
        if (s->elem_counts)
            check(ctx, sdbfx_set_add(ctx, dev, s->hashname, s->orig_file_size, s->bf_count, 0, s->dd_block_size,
                                     s->max_elem, s->buffer, s->elem_counts));
        else
            check(ctx, sdbfx_set_add(ctx, dev, s->hashname, s->orig_file_size, s->bf_count, s->last_count, 0,
                                     s->max_elem, s->buffer, nullptr));
    }
    // fast mode folds every item (sdbf/sdbf_set.cc:243-248), later ones too
    if (fast_)
        check(ctx, sdbfx_set_fold(ctx, dev));
    return dev;
}

static void put_line(std::ostream &out, const char *a, const char *b, char sep, int32_t score)
{
    out << a << sep << b;
    if (score != -1)
        out << sep << std::setw(3) << score << std::endl;
    else
        out << sep << score << std::endl;
}

// sdbf/sdbf_set.cc:278-308 (lines in i<j order)
std::string sdbf_set::compare_all_quiet(int32_t threshold, int32_t thread_count, bool fast)
{
    (void)thread_count;
    if (fast)
        fast_ = true;
    std::stringstream out;
    out.fill('0');
    const size_t n = items.size();
    if (n < 2)
        return out.str();
    std::vector<int32_t> scores(n * (n - 1) / 2);
    {
        std::lock_guard<std::mutex> g(sdbf::config->lock);
        sdbfx_set *d = device_set();
        check(sdbf::config->ctx(), sdbfx_compare_all(sdbf::config->ctx(), d, scores.data()));
    }
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++, k++)
            if (scores[k] >= threshold)
                put_line(out, items[i]->name(), items[j]->name(), sep, scores[k]);
    return out.str();
}

void sdbf_set::compare_all(int32_t threshold, bool fast)
{
    std::cout << compare_all_quiet(threshold, 0, fast);
}

// sdbf/sdbf_set.cc:360-394
std::string sdbf_set::compare_to_quiet(sdbf_set *other, int32_t threshold, uint32_t sample_size,
                                       int32_t thread_count, bool fast)
{
    (void)thread_count;
    (void)sample_size;
    if (fast)
        fast_ = other->fast_ = true;   // both sets folded (sdbf/sdbf_set.cc:322-333)
    std::stringstream out;
    out.fill('0');
    const size_t nq = items.size(), nt = other->items.size();
    if (!nq || !nt)
        return out.str();
    std::vector<int32_t> scores(nq * nt);
    {
        std::lock_guard<std::mutex> g(sdbf::config->lock);
        sdbfx_set *q = device_set();
        sdbfx_set *t = other->device_set();
        check(sdbf::config->ctx(), sdbfx_compare_to(sdbf::config->ctx(), q, t, scores.data()));
    }
    for (size_t i = 0; i < nq; i++)
        for (size_t j = 0; j < nt; j++)
            if (scores[i * nt + j] >= threshold)
                put_line(out, items[i]->name(), other->items[j]->name(), sep, scores[i * nt + j]);
    return out.str();
}

void sdbf_set::compare_to(sdbf_set *other, int32_t threshold, uint32_t sample_size, bool fast)
{
    std::cout << compare_to_quiet(other, threshold, sample_size, 0, fast);
}
