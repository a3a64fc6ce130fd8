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
#include <stdexcept>
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

// an index_info that asks for an index insert or search (sdbf_core.cc:110,
// 151-153, 165)
bool indexing(const index_info *info)
{
    return info && (info->index || (info->setlist && !info->setlist->empty()));
}

void check(sdbfx_ctx *ctx, int rc)
{
    if (rc != 0) {
        if (sdbf::config->warnings)
            std::cerr << "sdbf engine: " << sdbfx_last_error(ctx) << std::endl;
        throw -1;
    }
}

// LZ4 block format (the published "LZ4 Block Format Description"), which the
// reference's index files hold (bloom_filter.cc:176-200 through lz4.cc).
// A block is a run of sequences: a token (literal count in the high nibble,
// match length - 4 in the low one; 15 continues in following bytes, each
// added, until one below 255), the literals, a 2-byte little-endian offset
// back into the output and the match length's continuation bytes.  The last
// sequence holds literals only; the last 5 bytes are literals and no match
// starts within the last 12 bytes.
void lz4_length(std::vector<uint8_t> &o, size_t v)
{
    while (v >= 255) {
        o.push_back(255);
        v -= 255;
    }
    o.push_back((uint8_t)v);
}

// Greedy compressor: each 4-byte sequence is looked up in a table of the
// last position with the same hash.
std::vector<uint8_t> lz4_compress(const uint8_t *src, size_t n)
{
    std::vector<uint8_t> o;
    o.reserve(n / 32 + 64);
    std::vector<uint32_t> table(1u << 16, 0xFFFFFFFFu);
    size_t anchor = 0, i = 0;
    auto rd32 = [&](size_t p) {
        uint32_t v;
        memcpy(&v, src + p, 4);
        return v;
    };
    auto emit = [&](size_t lit_end, size_t off, size_t mlen) {
        const size_t ll = lit_end - anchor;
        uint8_t tok = (uint8_t)((ll >= 15 ? 15 : ll) << 4);
        if (mlen)
            tok |= (uint8_t)(mlen - 4 >= 15 ? 15 : mlen - 4);
        o.push_back(tok);
        if (ll >= 15)
            lz4_length(o, ll - 15);
        o.insert(o.end(), src + anchor, src + lit_end);
        if (mlen) {
            o.push_back((uint8_t)(off & 0xFF));
            o.push_back((uint8_t)(off >> 8));
            if (mlen - 4 >= 15)
                lz4_length(o, mlen - 4 - 15);
        }
    };
    if (n > 12) {
        const size_t limit = n - 12, mlimit = n - 5;
        while (i < limit) {
            const uint32_t v = rd32(i);
            const uint32_t h = (v * 2654435761u) >> 16;
            const uint32_t c = table[h];
            table[h] = (uint32_t)i;
            if (c != 0xFFFFFFFFu && i - c <= 65535 && rd32(c) == v) {
                size_t m = 4;
                while (i + m < mlimit && src[c + m] == src[i + m])
                    m++;
                emit(i, i - c, m);
                i += m;
                anchor = i;
            } else {
                i++;
            }
        }
    }
    emit(n, 0, 0);
    return o;
}

// 0 when src[0, srclen) decodes to exactly dstlen bytes, else -1
int lz4_decompress(const uint8_t *src, size_t srclen, uint8_t *dst, size_t dstlen)
{
    size_t ip = 0, op = 0;
    auto more = [&](size_t &v) {
        uint8_t b;
        do {
            if (ip >= srclen)
                return false;
            b = src[ip++];
            v += b;
        } while (b == 255);
        return true;
    };
    while (ip < srclen) {
        const uint8_t tok = src[ip++];
        size_t ll = tok >> 4;
        if (ll == 15 && !more(ll))
            return -1;
        if (ll > srclen - ip || ll > dstlen - op)
            return -1;
        memcpy(dst + op, src + ip, ll);
        ip += ll;
        op += ll;
        if (ip == srclen)
            break;
        if (srclen - ip < 2)
            return -1;
        const size_t off = (size_t)src[ip] | (size_t)src[ip + 1] << 8;
        ip += 2;
        size_t ml = tok & 15;
        if (ml == 15 && !more(ml))
            return -1;
        ml += 4;
        if (off == 0 || off > op || ml > dstlen - op)
            return -1;
        for (size_t k = 0; k < ml; k++)
            dst[op + k] = dst[op + k - off];
        op += ml;
    }
    return op == dstlen ? 0 : -1;
}

}  // namespace

// ----------------------------------------------------------------- bloom_filter
// sdbf/bloom_filter.cc:20-33, 44-64, 125-144, 263-359

const uint32_t bloom_filter::BITS[] = {0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80};

// BIT_MASKS_32[k] = 2^k - 1 for k >= 8, the low masks for k < 8
const uint32_t bloom_filter::BIT_MASKS_32[] = {
    0x01,       0x03,       0x07,       0x0F,       0x1F,       0x3F,       0x7F,       0xFF,
    0x01FF,     0x03FF,     0x07FF,     0x0FFF,     0x1FFF,     0x3FFF,     0x7FFF,     0xFFFF,
    0x01FFFF,   0x03FFFF,   0x07FFFF,   0x0FFFFF,   0x1FFFFF,   0x3FFFFF,   0x7FFFFF,   0xFFFFFF,
    0x01FFFFFF, 0x03FFFFFF, 0x07FFFFFF, 0x0FFFFFFF, 0x1FFFFFFF, 0x3FFFFFFF, 0x7FFFFFFF, 0xFFFFFFFF};

static uint16_t log_size_of(uint64_t size)
{
    uint16_t n = 0;
    for (uint64_t t = size; t; t >>= 1)
        n++;
    return n;
}

bloom_filter::bloom_filter(uint64_t size, uint16_t hash_count_, uint64_t max_elem_, double max_fp_)
    : hamming(0), hamminglg(0), bf_size(size), max_elem(max_elem_), hash_count(hash_count_), max_fp(max_fp_)
{
    if (!size || (size & (size - 1)) || size < 64 || log_size_of(size) + 1 >= 32)
        throw -1;
    bit_mask = BIT_MASKS_32[log_size_of(size) + 1];
    bf = (uint8_t *)calloc(size, 1);
}

bloom_filter::bloom_filter(uint8_t *data, uint64_t size, int id, int bf_elem_ct, uint16_t hamming_)
    : hamming(hamming_), hamminglg(0), bf_size(size), max_elem(192), hash_count(5)
{
    bit_mask = 2047;
    bf = (uint8_t *)malloc(size);
    memcpy(bf, data, size);
    bf_elem_count = (uint64_t)bf_elem_ct;
    bl_id = id;
}

// sdbf/bloom_filter.cc:77-114: "sdbf-idx:<size>:<elements>:<hash count>:
// <bit mask>:<compressed size>:<name>" on one line, then the compressed bits
bloom_filter::bloom_filter(std::string indexfilename) : hamming(0), hamminglg(0), bf_size(0), max_elem(0), hash_count(5)
{
    std::ifstream ifs(indexfilename.c_str(), std::ios::in | std::ios::binary);
    if (!ifs.is_open())
        throw -1;
    std::string f[6];
    for (int k = 0; k < 6; k++)
        std::getline(ifs, f[k], ':');
    std::getline(ifs, setname);
    uint64_t comp = 0;
    try {
        bf_size = std::stoull(f[1]);
        bf_elem_count = std::stoull(f[2]);
        hash_count = (uint16_t)std::stoul(f[3]);
        bit_mask = std::stoull(f[4]);
        comp = std::stoull(f[5]);
    } catch (...) {
        throw -2;
    }
    if (f[0] != "sdbf-idx" || !bf_size || (bf_size & (bf_size - 1)))
        throw -2;
    std::vector<uint8_t> c(comp);
    ifs.read((char *)c.data(), (std::streamsize)comp);
    if ((uint64_t)ifs.gcount() != comp)
        throw -2;
    bf = (uint8_t *)malloc(bf_size);
    if (lz4_decompress(c.data(), comp, bf, bf_size) != 0) {
        free(bf);
        throw -2;
    }
}

// sdbf/bloom_filter.cc:208-228
int32_t bloom_filter::write_out(std::string filename)
{
    const std::vector<uint8_t> c = lz4_compress(bf, bf_size);
    if (c.empty())
        return -1;
    std::ofstream os(filename.c_str(), std::ios::out | std::ios::binary);
    if (!os.is_open())
        return -2;
    os << "sdbf-idx:" << bf_size << ":" << bf_elem_count << ":" << hash_count << ":" << bit_mask << ":" << c.size()
       << ":" << setname << std::endl;
    os.write((const char *)c.data(), (std::streamsize)c.size());
    return os.good() ? 0 : -2;
}

bloom_filter::~bloom_filter() { free(bf); }

bool bloom_filter::query_and_set(uint32_t *sha1, bool mode_set)
{
    uint32_t bit_cnt = 0;
    for (uint32_t i = 0; i < hash_count; i++) {
        const uint64_t pos = sha1[i] & bit_mask;
        const uint8_t bit = (uint8_t)BITS[pos & 7];
        if (bf[pos >> 3] & bit)
            bit_cnt++;
        else if (mode_set)
            bf[pos >> 3] |= bit;
        else
            return false;
    }
    if (!mode_set)
        return bit_cnt == hash_count;
    if (bit_cnt < hash_count) {
        bf_elem_count++;
        return true;
    }
    return false;
}

bool bloom_filter::insert_sha1(uint32_t *sha1) { return query_and_set(sha1, true); }
bool bloom_filter::query_sha1(uint32_t *sha1) { return query_and_set(sha1, false); }
uint64_t bloom_filter::elem_count() { return bf_elem_count; }

// the reference returns -1 (no estimate, sdbf/bloom_filter.cc:163)
double bloom_filter::est_fp_rate() { return -1.0; }

double bloom_filter::bits_per_elem() { return (double)(bf_size << 3) / (double)bf_elem_count; }
std::string bloom_filter::name() const { return setname; }
void bloom_filter::set_name(std::string n) { setname = n; }
int bloom_filter::bloom_id() { return bl_id; }
void bloom_filter::set_bloom_id(int id) { bl_id = id; }

// halves the filter `times` times by OR-ing its upper half onto its lower
// half; as in sdbf/bloom_filter.cc:263-289 a halving that reaches 32 words
// (256 bytes) ends the folding
void bloom_filter::fold(uint32_t times)
{
    uint64_t rsize = bf_size / 8;
    uint64_t *w = (uint64_t *)bf;
    for (uint32_t i = 0; i < times; i++) {
        for (uint64_t j = 0; j < rsize / 2; j++)
            w[j] |= w[j + rsize / 2];
        rsize /= 2;
        if (rsize == 32)
            break;
    }
    bf_size = rsize * 8;
    bit_mask = BIT_MASKS_32[log_size_of(bf_size) + 1];
    uint8_t *nb = (uint8_t *)malloc(bf_size);
    memcpy(nb, bf, bf_size);
    free(bf);
    bf = nb;
}

int bloom_filter::add(bloom_filter *other)
{
    if (bf_size != other->bf_size)
        return 1;
    for (uint64_t j = 0; j < bf_size; j++)
        bf[j] |= other->bf[j];
    return 0;
}

void bloom_filter::compute_hamming()
{
    hamming = 0;
    hamminglg = 0;
    for (uint64_t j = 0; j < bf_size; j++)
        hamminglg += (uint32_t)__builtin_popcount(bf[j]);
}

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

// The digests of inputs [0, n) in a new device set.  With an index_info
// that asks for indexing (sdbf_core.cc:110-115, 157-190): info->index is
// copied to the device, gates / takes the inserts of the inputs in order, and
// is copied back; the indexes of info->setlist are looked up by block-mode
// inputs, and each block with more than _FP_THRESHOLD (4) hits in an index
// gives a print_indexes line "<name> [<block>] |<set name>|<hits>" (sdbf_core.
// cc:458-472), unless search_first or search_deep is set (:184-187).
sdbfx_set *sdbf::digest_inputs(sdbfx_ctx *ctx, const std::vector<const char *> &names,
                               const std::vector<const uint8_t *> &bufs, const std::vector<uint64_t> &lens,
                               const std::vector<uint32_t> &dd, std::vector<int32_t> &status, index_info *info,
                               std::vector<std::string> &results)
{
    const uint32_t n = (uint32_t)bufs.size();
    status.assign(n, 0);
    results.assign(n, std::string());
    sdbfx_set *set = nullptr;
    check(ctx, sdbfx_set_create(ctx, &set));
    if (!indexing(info)) {
        const int rc = sdbfx_set_digest(ctx, set, n, bufs.data(), lens.data(), names.data(), dd.data(), status.data());
        if (rc != 0) {
            sdbfx_set_destroy(set);
            check(ctx, rc);
        }
        return set;
    }
    sdbfx_index *ix = nullptr;
    std::vector<sdbfx_index *> search;
    std::vector<sdbf_set *> sets;
    auto release = [&]() {
        sdbfx_index_destroy(ix);
        for (sdbfx_index *p : search)
            sdbfx_index_destroy(p);
    };
    auto to_device = [&](bloom_filter *bf) {
        sdbfx_index *d = nullptr;
        if (sdbfx_index_create(ctx, bf->bf_size, &d) != 0 ||
            sdbfx_index_write(ctx, d, bf->bf, bf->bf_elem_count) != 0) {
            sdbfx_index_destroy(d);
            return (sdbfx_index *)nullptr;
        }
        return d;
    };
    int rc = 0;
    if (info->index && !(ix = to_device(info->index)))
        rc = -1;
    if (info->setlist)
        for (sdbf_set *ss : *info->setlist) {
            if (rc == 0 && (!ss || !ss->index))
                rc = -1;
            if (rc == 0) {
                sdbfx_index *d = to_device(ss->index);
                if (!d)
                    rc = -1;
                else
                    search.push_back(d), sets.push_back(ss);
            }
        }
    uint64_t nblk = 0;
    for (uint32_t i = 0; i < n; i++)
        if (dd[i] && lens[i] >= 512)
            nblk += lens[i] / dd[i] + (lens[i] % dd[i] >= 512 ? 1 : 0);
    const uint32_t ns = (uint32_t)search.size();
    std::vector<uint32_t> matches(std::max<uint64_t>(nblk * ns, 1));
    if (rc == 0)
        rc = sdbfx_set_digest_indexed(ctx, set, n, bufs.data(), lens.data(), names.data(), dd.data(), ix,
                                      (const sdbfx_index *const *)search.data(), ns, matches.data(), nblk,
                                      status.data());
    if (rc == 0 && ix) {
        uint64_t e = 0;
        rc = sdbfx_index_read(ctx, ix, info->index->bf, &e);
        info->index->bf_elem_count = e;
    }
    release();
    if (rc != 0) {
        sdbfx_set_destroy(set);
        check(ctx, rc);
    }
    if (ns && !info->search_first && !info->search_deep) {
        uint64_t b = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!dd[i] || lens[i] < 512)
                continue;
            const uint64_t cnt = lens[i] / dd[i] + (lens[i] % dd[i] >= 512 ? 1 : 0);
            std::stringstream build;
            for (uint64_t k = 0; k < cnt; k++, b++)
                for (uint32_t m = 0; m < ns; m++)
                    if (matches[b * ns + m] > 4)
                        build << names[i] << " [" << k << "] |" << sets[m]->name() << "|" << matches[b * ns + m]
                              << std::endl;
            results[i] = build.str();
        }
    }
    return set;
}

void sdbf::hash_buffer(const char *name, const uint8_t *buf, uint64_t len, uint32_t dd, index_info *info)
{
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    const char *nm = name ? name : "";
    std::vector<int32_t> status;
    std::vector<std::string> res;
    sdbfx_set *set = digest_inputs(ctx, {nm}, {buf}, {len}, {dd}, status, info, res);
    if (status[0] != 0) {
        sdbfx_set_destroy(set);
        throw (int)status[0];
    }
    from_device(set, 0, nm);
    index_results = res[0];
    sdbfx_set_destroy(set);
}

void sdbf::hash_batch(const std::vector<std::string> &names, const std::vector<const uint8_t *> &bufs,
                      const std::vector<uint64_t> &lens, uint32_t dd, std::vector<sdbf *> &out,
                      const std::vector<uint64_t> *orig_sizes, index_info *info)
{
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    const uint32_t n = (uint32_t)bufs.size();
    std::vector<const char *> nm(n);
    std::vector<uint32_t> ddv(n, dd);
    for (uint32_t i = 0; i < n; i++)
        nm[i] = names[i].c_str();
    std::vector<int32_t> status;
    std::vector<std::string> res;
    sdbfx_set *set = digest_inputs(ctx, nm, bufs, lens, ddv, status, info, res);
    out.assign(n, nullptr);
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (status[i] != 0)
            continue;
        sdbf *s = new sdbf();
        s->from_device(set, k++, nm[i]);
        if (orig_sizes)
            s->orig_file_size = (*orig_sizes)[i];
        s->index_results = res[i];
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
    hash_buffer(filename, data.data(), data.size(), dd, nullptr);
}

// sdbf/sdbf_class.cc:71-102: reads msize bytes; fewer than 512 throw -3; the
// digest covers all msize bytes (zero filled past the read)
sdbf::sdbf(const char *name, std::istream *ifs, uint32_t dd, uint64_t msize, index_info *info)
{
    std::vector<uint8_t> data(msize, 0);
    ifs->read((char *)data.data(), (std::streamsize)msize);
    const uint64_t got = (uint64_t)ifs->gcount();
    if (got < 512)
        throw -3;
    hash_buffer(name, data.data(), msize, dd, info);
    orig_file_size = got;
}

// sdbf/sdbf_class.cc:115-136
sdbf::sdbf(const char *name, char *str, uint32_t dd, uint64_t length, index_info *info)
{
    if (length < 512)
        throw -3;
    hash_buffer(name, (const uint8_t *)str, length, dd, info);
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
    delete big_filters;
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

// uploads one digest to a device set, folded on the device when the host
// object is in fast mode (its host filters are folded already; folding is
// idempotent, and it sets the device's Hamming weights to the folded ones)
static int add_to_device(sdbfx_ctx *ctx, sdbfx_set *set, sdbf *s)
{
    const uint32_t idx = sdbfx_set_size(set);
    int rc = sdbfx_set_add(ctx, set, s->name(), s->input_size(), s->filter_count(), s->last_count_(),
                           s->dd_block_size_(), s->max_elem, s->buffer, s->elem_counts_());
    if (rc == 0 && s->fast_mode())
        rc = sdbfx_set_fold_digest(ctx, set, idx);
    return rc;
}

// sdbf/sdbf_class.cc:266-271 -> sdbf_score (sdbf/sdbf_core.cc:331-389),
// `sample` included
int32_t sdbf::compare(sdbf *other, uint32_t sample)
{
    std::lock_guard<std::mutex> g(config->lock);
    sdbfx_ctx *ctx = config->ctx();
    sdbfx_set *q = nullptr, *t = nullptr;
    check(ctx, sdbfx_set_create(ctx, &q));
    int rc = sdbfx_set_create(ctx, &t);
    if (rc == 0)
        rc = add_to_device(ctx, q, this);
    if (rc == 0)
        rc = add_to_device(ctx, t, other);
    int32_t score = -1;
    if (rc == 0)
        rc = sdbfx_compare_to_sampled(ctx, q, t, sample, &score);
    sdbfx_set_destroy(q);
    sdbfx_set_destroy(t);
    check(ctx, rc);
    return score;
}

// sdbf/sdbf_class.cc:426-439 with bloom_filter::fold(2) and compute_hamming
// (sdbf/bloom_filter.cc:263-289, 453-461): each filter's 32 64-bit words
// folded twice (w[j] |= w[j + 16], j < 16; then w[j] |= w[j + 8], j < 8), the
// 64 folded bytes written over the filter's first 64 bytes, its Hamming
// weight their popcount
void sdbf::fast()
{
    for (uint32_t i = 0; i < bf_count; i++) {
        uint64_t w[32];
        memcpy(w, buffer + (size_t)i * bf_size, 256);
        for (int j = 0; j < 16; j++)
            w[j] |= w[j + 16];
        for (int j = 0; j < 8; j++)
            w[j] |= w[j + 8];
        memcpy(buffer + (size_t)i * bf_size, w, 64);
        uint32_t h = 0;
        for (int j = 0; j < 8; j++)
            h += (uint32_t)__builtin_popcountll(w[j]);
        hamming[i] = (uint16_t)h;
    }
    fastmode = true;
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

sdbf_set::sdbf_set() : bf_vector(new std::vector<bloom_filter *>()) {}

// sdbf/sdbf_set.cc:40-46
sdbf_set::sdbf_set(bloom_filter *idx) : index(idx), bf_vector(new std::vector<bloom_filter *>()) {}

// sdbf/sdbf_set.cc:51-74
sdbf_set::sdbf_set(const char *fname) : bf_vector(new std::vector<bloom_filter *>())
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
    vector_init();
}

// sdbf/sdbf_set.cc:80-86: the set's digests are not deleted, bf_vector is
sdbf_set::~sdbf_set()
{
    if (dev)
        sdbfx_set_destroy(dev);
    for (bloom_filter *b : *bf_vector)
        delete b;
    delete bf_vector;
}

// sdbf/sdbf_set.cc:409-420: a host copy of every filter of every digest
void sdbf_set::vector_init()
{
    for (uint32_t i = 0; i < items.size(); i++)
        for (uint32_t n = 0; n < items[i]->filter_count(); n++) {
            bloom_filter *b = new bloom_filter(items[i]->buffer + (size_t)n * 256, 256, (int)i,
                                               sdbf::get_elem_count(items[i], n), items[i]->hamming[n]);
            b->set_name(items[i]->name());
            bf_vector->push_back(b);
        }
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

// sdbf/sdbf_set.cc:401-404: the size of bf_vector (filled by vector_init,
// which the file constructor calls)
uint64_t sdbf_set::filter_count() { return bf_vector->size(); }

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

// sdbf/sdbf_set.cc:161-167: the items' print_indexes lines, in set order
std::string sdbf_set::index_results() const
{
    std::string r;
    for (const sdbf *it : items)
        r += it->get_index_results();
    return r;
}
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
    for (; dev_items < items.size(); dev_items++)
        check(ctx, add_to_device(ctx, dev, items[dev_items]));
    // items folded (sdbf::fast) after their upload
    for (uint32_t i = 0; i < items.size(); i++)
        if (items[i]->fast_mode() && !sdbfx_set_digest_folded(dev, i))
            check(ctx, sdbfx_set_fold_digest(ctx, dev, i));
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
    // sdbf/sdbf_set.cc:287-292: every item folded
    if (fast)
        for (sdbf *s : items)
            s->fast();
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
    // sdbf/sdbf_set.cc:367-376: fast mode folds this->items.at(i) for i below
    // other's size, then for i below this set's size -- so only this set is
    // folded, and a larger other set makes at() throw std::out_of_range
    // (inside an OpenMP region in the reference, which terminates)
    if (fast) {
        if (other->items.size() > items.size())
            throw std::out_of_range("sdbf_set::compare_to: fast mode folds this->items.at(i) for every i "
                                    "below the other set's size");
        for (sdbf *s : items)
            s->fast();
    }
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
        check(sdbf::config->ctx(), sdbfx_compare_to_sampled(sdbf::config->ctx(), q, t, sample_size, scores.data()));
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
