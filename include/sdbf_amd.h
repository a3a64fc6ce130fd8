// sdbf_amd.h -- C++ classes with the interface of the reference's libsdbf
// (sdbf/sdbf_class.h:23-121, sdbf/sdbf_set.h:15-96), implemented on the
// gfx950 engine behind include/sdbfx.h.  Callers written against the
// reference (sdhash-src, sdhash-server) recompile against this header.
//
// Differences from the reference, all outside the hot path:
//   - index_info (--index / --index-search) is accepted as a parameter but
//     must be NULL; sampling (`sample` > 0) is not supported and throws -1.
//   - fast mode (`fast` = true) folds the set's device copy (sdbfx_set_fold),
//     as sdbf::fast folds each item; the host sdbf objects keep their
//     unfolded filters, so to_string after a fast comparison prints those.
//     Once a set has been compared in fast mode it stays folded, as in the
//     reference.
//   - sdbf_set::bf_vector (the per-filter bloom_filter copies of
//     vector_init, sdbf/sdbf_set.cc:409-420) is not materialised; the
//     device-resident set replaces it.
#ifndef SDBF_AMD_H
#define SDBF_AMD_H

#include <stdint.h>

#include <iostream>
#include <mutex>
#include <string>
#include <vector>

#include "sdbfx.h"

struct index_info;   // reference --index support: accepted only as NULL

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
    std::string get_index_results() const { return std::string(); }
    uint8_t *clone_filter(uint32_t position);
    uint32_t filter_count();

    static sdbf_conf *config;
    static int32_t get_elem_count(sdbf *mine, uint64_t index);

    // Digests many buffers in one device pass (the batch form of the
    // buffer constructor).  Entries of `out` are NULL for inputs that are
    // too small (the constructor's -3).
    static void hash_batch(const std::vector<std::string> &names, const std::vector<const uint8_t *> &bufs,
                           const std::vector<uint64_t> &lens, uint32_t dd_block_size, std::vector<sdbf *> &out);

public:
    uint8_t *buffer = nullptr;     // bf_count * 256 filter bytes
    uint16_t *hamming = nullptr;   // Hamming weight of every filter
    uint32_t max_elem = 0;

private:
    sdbf() = default;
    void from_device(sdbfx_set *set, uint32_t idx, const char *name);
    void hash_buffer(const char *name, const uint8_t *buf, uint64_t len, uint32_t dd);

    char *hashname = nullptr;
    uint32_t bf_count = 0;
    uint32_t bf_size = 256;
    uint32_t hash_count = 5;
    uint32_t mask = 0x7FF;
    uint32_t last_count = 0;
    uint16_t *elem_counts = nullptr;
    uint32_t dd_block_size = 0;
    uint64_t orig_file_size = 0;
};

class sdbf_set {
    friend std::ostream &operator<<(std::ostream &os, const sdbf_set &s);
    friend std::ostream &operator<<(std::ostream &os, const sdbf_set *s);

public:
    sdbf_set();
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
    std::string index_results() const { return std::string(); }
    std::string name() const;
    void set_name(std::string name);
    void set_separator(char sep);
    void vector_init() {}

private:
    sdbfx_set *device_set();   // the items uploaded to the device (cached)
    std::vector<sdbf *> items;
    std::string setname = "default";
    std::mutex add_hash_mutex;
    char sep = '|';
    sdbfx_set *dev = nullptr;
    size_t dev_items = 0;
    bool fast_ = false;        // compared in fast mode: the device copy is folded
};

std::ostream &operator<<(std::ostream &os, const sdbf &s);
std::ostream &operator<<(std::ostream &os, const sdbf *s);
std::ostream &operator<<(std::ostream &os, const sdbf_set &s);
std::ostream &operator<<(std::ostream &os, const sdbf_set *s);

#endif
