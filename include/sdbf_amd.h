// sdbf_amd.h -- C++ classes with the interface of the reference's libsdbf
// (sdbf/sdbf_class.h:23-121, sdbf/sdbf_set.h:15-96, sdbf/bloom_filter.h:
// 16-101), implemented on the gfx950 engine behind include/sdbfx.h.  Callers
// written against the reference (sdhash-src, sdhash-server) recompile
// against this header (sdbf_defines.h includes it under the reference's
// header name).
//
// Differences from the reference, all outside the hot path:
//   - index_info (--index / --index-search): the indexes are used on the
//     device; an info->index is copied to the device for each digest call
//     and its bits and element count copied back afterwards, so its content
//     is the reference's after every constructor or hash_batch call.  A set
//     list entry without an index throws -1 (the reference dereferences it).
//     bloom_filter::write_out writes the reference's .idx header and a valid
//     LZ4 block of the bits (any LZ4 decoder, the reference's included,
//     reads it back), which need not be byte-identical with lz4's own
//     compressor output.
//   - sampling (compare(other, sample), compare_to(..., sample_size, ...))
//     gives the reference's scores: a calling digest with more than `sample`
//     filters is scored on its first `sample` filters; the reference's
//     random offsets are unreachable (sdbf/sdbf_core.cc:342-374).
//   - sdbf::big_filters holds no filters: the 16 KiB de-duplication filters
//     of stream digesting live in device memory for the duration of one
//     digest only (the reference keeps them for --large output, which needs
//     protobuf).
//   - sdbf_set::bf_vector is filled by vector_init() with a host copy of
//     every filter (sdbf/sdbf_set.cc:409-420); the comparisons themselves
//     use the device-resident copy of the set.
#ifndef SDBF_AMD_H
#define SDBF_AMD_H

#include <stdint.h>

#include <iostream>
#include <mutex>
#include <string>
#include <vector>

#include "sdbfx.h"

class bloom_filter;
class sdbf_set;

// --index / --index-search state (sdbf/index_info.h:8-15).  The reference
// CLI passes one to every constructor; with no index (index == NULL and no
// set list to search) it changes nothing.
struct index_info {
    bloom_filter *index;
    std::vector<bloom_filter *> *indexlist;
    std::vector<sdbf_set *> *setlist;
    bool search_deep;
    bool search_first;
    bool basename;
};

// A Bloom filter (sdbf/bloom_filter.h:16-101): the reference's container for
// sdbf_set::index and sdbf_set::bf_vector entries, with its public fields.
// Host-side; the digest and comparison kernels hold their filters on the
// device.
class bloom_filter {
public:
    // empty filter of `size` bytes (a power of 2, >= 64; otherwise throws -1)
    bloom_filter(uint64_t size, uint16_t hash_count, uint64_t max_elem, double max_fp);
    // copy of a 256-byte sdbf filter (sdbf/bloom_filter.cc:125-144)
    bloom_filter(uint8_t *data, uint64_t size, int id, int bf_elem_ct, uint16_t hamming);
    // reads an index file written by write_out (sdbf/bloom_filter.cc:77-114):
    // throws -1 when it cannot be read, -2 when it does not decompress
    bloom_filter(std::string indexfilename);
    // writes the index file (sdbf/bloom_filter.cc:208-228): -1 when the
    // compression fails, -2 when the file cannot be written
    int32_t write_out(std::string filename);
    ~bloom_filter();
    bool insert_sha1(uint32_t *sha1);
    bool query_sha1(uint32_t *sha1);
    uint64_t elem_count();
    double est_fp_rate();
    double bits_per_elem();
    std::string name() const;
    void set_name(std::string name);
    void fold(uint32_t times);
    int add(bloom_filter *other);
    int bloom_id();
    void set_bloom_id(int id);
    void compute_hamming();

    static const uint32_t BIT_MASKS_32[];
    static const uint32_t BITS[];

    uint8_t *bf;
    uint16_t hamming;
    uint32_t hamminglg;
    uint64_t bf_size;
    uint64_t bit_mask;
    uint64_t max_elem;
    uint16_t hash_count;

private:
    friend class sdbf;
    bool query_and_set(uint32_t *sha1, bool mode_set);
    double max_fp = 0;
    uint64_t bf_elem_count = 0;
    std::string setname;
    int bl_id = 0;
};

// Engine-wide configuration, the counterpart of sdbf_conf
// (sdbf/sdbf_conf.h:16-58): the constants the path uses and the device
// context shared by all digests of the process.
class sdbf_conf {
public:
    sdbf_conf(uint32_t thread_cnt, uint32_t warnings, uint32_t max_elem_ct, uint32_t max_elem_ct_dd);
    ~sdbf_conf();
    uint32_t thread_cnt;
    uint32_t entr_win_size = 64;
    uint32_t bf_size = 256;
    uint32_t pop_win_size = 64;
    uint32_t block_size = 4096;
    uint32_t max_elem;
    uint32_t max_elem_dd;
    uint32_t warnings;
    uint32_t threshold = 16;
    int device = 0;
    sdbfx_ctx *ctx();          // opened on first use
    std::mutex lock;           // serialises use of the shared context
private:
    sdbfx_ctx *ctx_ = nullptr;
};

class sdbf {
    friend std::ostream &operator<<(std::ostream &os, const sdbf &s);
    friend std::ostream &operator<<(std::ostream &os, const sdbf *s);
    friend class sdbf_set;

public:
    sdbf(const std::string &str);                                    // parse text form
    sdbf(const char *filename, uint32_t dd_block_size);              // hash a file
    sdbf(const char *name, std::istream *ifs, uint32_t dd_block_size, uint64_t msize, index_info *info);
    sdbf(const char *name, char *str, uint32_t dd_block_size, uint64_t length, index_info *info);
    ~sdbf();

    const char *name();
    uint64_t size();
    uint64_t input_size();
    int32_t compare(sdbf *other, uint32_t sample);
    std::string to_string() const;
    std::string get_index_results() const { return index_results; }
    uint8_t *clone_filter(uint32_t position);
    uint32_t filter_count();
    // fast mode (sdbf/sdbf_class.cc:426-439): folds every filter to 64 bytes
    // in place; comparisons with this digest as the reference side then
    // score on the folded filters
    void fast();

    static sdbf_conf *config;
    static int32_t get_elem_count(sdbf *mine, uint64_t index);

    // Digests many buffers in one device pass (the batch form of the
    // buffer constructor).  Entries of `out` are NULL for inputs that are
    // too small (the constructor's -3).  orig_sizes (optional) gives each
    // digest's input_size() when it differs from the digested length, as
    // for the stream constructor, whose digest covers msize bytes of which
    // it read fewer.
    // With an index_info (info), the inputs are digested in the order given
    // against its indexes, as the constructors do one after another.
    static void hash_batch(const std::vector<std::string> &names, const std::vector<const uint8_t *> &bufs,
                           const std::vector<uint64_t> &lens, uint32_t dd_block_size, std::vector<sdbf *> &out,
                           const std::vector<uint64_t> *orig_sizes = nullptr, index_info *info = nullptr);

public:
    uint8_t *buffer = nullptr;     // bf_count * 256 filter bytes
    uint16_t *hamming = nullptr;   // Hamming weight of every filter
    uint32_t max_elem = 0;
    std::vector<bloom_filter *> *big_filters = new std::vector<bloom_filter *>();   // always empty (see above)

    // fields the reference keeps private, for the engine's own code
    uint32_t last_count_() const { return last_count; }
    uint32_t dd_block_size_() const { return dd_block_size; }
    const uint16_t *elem_counts_() const { return elem_counts; }
    bool fast_mode() const { return fastmode; }

private:
    sdbf() = default;
    void from_device(sdbfx_set *set, uint32_t idx, const char *name);
    void hash_buffer(const char *name, const uint8_t *buf, uint64_t len, uint32_t dd, index_info *info);
    // one device digest call (with info's indexes when it asks for them);
    // results[k] gets the print_indexes lines of the k-th input
    static sdbfx_set *digest_inputs(sdbfx_ctx *ctx, const std::vector<const char *> &names,
                                    const std::vector<const uint8_t *> &bufs, const std::vector<uint64_t> &lens,
                                    const std::vector<uint32_t> &dd, std::vector<int32_t> &status, index_info *info,
                                    std::vector<std::string> &results);

    char *hashname = nullptr;
    uint32_t bf_count = 0;
    uint32_t bf_size = 256;
    uint32_t hash_count = 5;
    uint32_t mask = 0x7FF;
    uint32_t last_count = 0;
    uint16_t *elem_counts = nullptr;
    uint32_t dd_block_size = 0;
    uint64_t orig_file_size = 0;
    bool fastmode = false;
    std::string index_results;   // print_indexes lines of --index-search
};

class sdbf_set {
    friend std::ostream &operator<<(std::ostream &os, const sdbf_set &s);
    friend std::ostream &operator<<(std::ostream &os, const sdbf_set *s);

public:
    sdbf_set();
    sdbf_set(bloom_filter *index);  // empty set with an index (sdbf/sdbf_set.cc:40-46)
    sdbf_set(const char *fname);   // load a file of sdbf lines
    ~sdbf_set();
    sdbf *at(uint32_t pos);
    void add(sdbf *hash);
    void add(sdbf_set *hashset);
    uint64_t size();
    uint64_t input_size();
    uint64_t filter_count();
    int empty();
    void compare_all(int32_t threshold, bool fast);
    std::string compare_all_quiet(int32_t threshold, int32_t thread_count, bool fast);
    void compare_to(sdbf_set *other, int32_t threshold, uint32_t sample_size, bool fast);
    std::string compare_to_quiet(sdbf_set *other, int32_t threshold, uint32_t sample_size, int32_t thread_count,
                                 bool fast);
    std::string to_string() const;
    std::string index_results() const;
    std::string name() const;
    void set_name(std::string name);
    void set_separator(char sep);
    void vector_init();

public:
    bloom_filter *index = nullptr;
    std::vector<bloom_filter *> *bf_vector = nullptr;

private:
    sdbfx_set *device_set();   // the items uploaded to the device (cached)
    std::vector<sdbf *> items;
    std::string setname = "default";
    std::mutex add_hash_mutex;
    char sep = '|';
    sdbfx_set *dev = nullptr;
    size_t dev_items = 0;
};

std::ostream &operator<<(std::ostream &os, const sdbf &s);
std::ostream &operator<<(std::ostream &os, const sdbf *s);
std::ostream &operator<<(std::ostream &os, const sdbf_set &s);
std::ostream &operator<<(std::ostream &os, const sdbf_set *s);

#endif
