// A caller written against the reference's libsdbf interface, using its call
// sites as they stand (with this engine's headers under the reference's
// names): the CLI's hashing loops (sdhash-src/sdhash_threads.cc:153-373) and
// comparisons (sdhash-src/sdhash.cc:316-392), and the server's set handling
// (sdhash-server/src/sdhash-srv-main.cc:224-402, 626-633).
//
//   caller_compat hash <dd KiB> <segment MiB> files...   digests, as text
//   caller_compat stdin <dd KiB> <segment MiB> name       digests of stdin
//   caller_compat all <threshold> <fast 0|1> set.sdbf      compare_all_quiet
//   caller_compat two <threshold> <sample> <fast 0|1> q.sdbf t.sdbf
//   caller_compat api                                      member checks
//   caller_compat idxw <size> <n> <seed> <file.idx>        index of n inserts, written out
//   caller_compat idxr <file.idx> <bits.bin>               index file read back
#include <sys/stat.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "bloom_filter.h"
#include "index_info.h"
#include "sdbf_class.h"
#include "sdbf_defines.h"
#include "sdbf_set.h"

static std::vector<sdbf_set *> setlist;

// sdhash-src/sdhash.cc:234-240
static index_info *make_info()
{
    index_info *info = (index_info *)malloc(sizeof(index_info));
    info->index = NULL;
    info->indexlist = NULL;
    info->setlist = &setlist;
    info->search_deep = false;
    info->search_first = false;
    info->basename = false;
    return info;
}

// sdbf_hash_files (sdhash-src/sdhash_threads.cc:241-297), one thread
static void hash_files(char **filenames, uint32_t file_count, sdbf_set *addto, index_info *info)
{
    std::ifstream *is = new std::ifstream();
    for (uint32_t i = 0; i < file_count; i++) {
        struct stat file_stat;
        if (stat(filenames[i], &file_stat))
            continue;
        is->open(filenames[i], std::ios::binary);
        try {
            class sdbf *sdbfm = new sdbf(filenames[i], is, 0, file_stat.st_size, info);
            addto->add(sdbfm);
        } catch (int e) {
            if (e == -3)
                std::cerr << "Input file too small for processing: " << filenames[i] << std::endl;
        }
        is->close();
    }
    delete is;
}

// sdbf_hash_files_dd (sdhash-src/sdhash_threads.cc:300-373), with its
// segment rules: no segments when the file is at most 512 bytes over one
// segment; a last segment of at most 512 bytes is merged into the one before
// it, whose length then counts the tail twice (the stream read stops at the
// file's end, so the digest covers the tail's length again in zeros)
static void hash_files_dd(char **filenames, uint32_t file_count, uint32_t dd_block_size, uint64_t chunk_size,
                          sdbf_set *addto, index_info *info)
{
    std::ifstream *is = new std::ifstream();
    for (uint32_t i = 0; i < file_count; i++) {
        struct stat file_stat;
        if (stat(filenames[i], &file_stat))
            continue;
        const uint64_t filesize = (uint64_t)file_stat.st_size;
        is->open(filenames[i], std::ios::binary);
        if (filesize > chunk_size && chunk_size > 0) {
            uint64_t chunks = filesize / chunk_size;
            bool tail = false;
            if (filesize - chunk_size <= 512)
                chunks = 0;
            if (filesize - chunks * chunk_size <= 512) {
                chunks--;
                tail = true;
            }
            for (uint64_t j = 0; j <= chunks; j++) {
                std::stringstream namestr;
                namestr << filenames[i];
                if (j > 0 || chunks > 0) {
                    namestr.fill('0');
                    namestr << "." << std::setw(4) << j * chunk_size / MB << "M";
                }
                std::string myname = namestr.str();
                char *fname = (char *)malloc(myname.length() + 1);
                strncpy(fname, myname.c_str(), myname.length() + 1);
                uint64_t csize = chunk_size;
                if (j == chunks) {
                    csize = filesize - chunks * chunk_size;
                    if (tail)
                        csize += filesize - (chunks + 1) * chunk_size;
                }
                try {
                    class sdbf *sdbfm = new sdbf(fname, is, dd_block_size, csize, info);
                    addto->add(sdbfm);
                } catch (int e) {
                    if (e == -2)
                        exit(-2);
                }
            }
        } else {
            try {
                class sdbf *sdbfm = new sdbf(filenames[i], is, dd_block_size, filesize, info);
                addto->add(sdbfm);
            } catch (int e) {
                if (e == -2)
                    exit(-2);
            }
        }
        is->close();
    }
    delete is;
}

// sdbf_hash_stdin (sdhash-src/sdhash_threads.cc:190-236)
static sdbf_set *hash_stdin(const char *stream, uint32_t dd_block_size, uint64_t segment_size, index_info *info)
{
    sdbf_set *myset = new sdbf_set();
    char *readbuf = (char *)malloc(segment_size);
    for (uint64_t j = 0;; j++) {
        memset(readbuf, 0, segment_size);
        size_t sz = fread(readbuf, 1, segment_size, stdin);
        std::stringstream namestr;
        namestr.fill('0');
        namestr << stream << "." << std::setw(4) << j * segment_size / MB << "M";
        std::string myname = namestr.str();
        char *fname = (char *)malloc(myname.length() + 1);
        strncpy(fname, myname.c_str(), myname.length() + 1);
        try {
            class sdbf *sdbfm = new sdbf(fname, readbuf, dd_block_size * KB, sz, info);
            myset->add(sdbfm);
        } catch (int e) {
            if (e == -2)
                exit(-2);
        }
        if (sz != segment_size)
            break;
    }
    free(readbuf);
    return myset;
}

// loadSet + add_set (sdhash-srv-main.cc:264-300, 626-633)
static sdbf_set *load_set(const std::string &filename)
{
    sdbf_set *tmp_set = new sdbf_set(filename.c_str());
    if (tmp_set->empty())
        return tmp_set;
    if (tmp_set->index != NULL)
        setlist.push_back(tmp_set);
    tmp_set->set_name(filename);
    return tmp_set;
}

int main(int argc, char **argv)
{
    if (argc < 2)
        return 2;
    const std::string mode = argv[1];
    index_info *info = make_info();
    if (mode == "hash" && argc >= 4) {
        sdbf_set *set1 = new sdbf_set();
        const uint32_t dd = (uint32_t)atoi(argv[2]);
        const uint64_t seg = (uint64_t)atoll(argv[3]) * MB;
        if (dd)
            hash_files_dd(argv + 4, (uint32_t)(argc - 4), dd * KB, seg, set1, info);
        else
            hash_files(argv + 4, (uint32_t)(argc - 4), set1, info);
        std::cout << set1;
        delete set1;
    } else if (mode == "stdin" && argc == 5) {
        sdbf_set *set1 = hash_stdin(argv[4], (uint32_t)atoi(argv[2]), (uint64_t)atoll(argv[3]) * MB, info);
        std::cout << set1;
        delete set1;
    } else if (mode == "all" && argc == 5) {
        sdbf_set *set1 = load_set(argv[4]);
        std::cout << set1->compare_all_quiet(atoi(argv[2]), 1, atoi(argv[3]) != 0);
        std::cerr << "filters " << set1->filter_count() << " input " << set1->input_size() << std::endl;
    } else if (mode == "two" && argc == 7) {
        sdbf_set *set1 = load_set(argv[5]);
        sdbf_set *set2 = load_set(argv[6]);
        try {
            std::cout << set1->compare_to_quiet(set2, atoi(argv[2]), (uint32_t)atoi(argv[3]), 1, atoi(argv[4]) != 0);
        } catch (const std::out_of_range &) {
            std::cerr << "out_of_range" << std::endl;
            return 3;
        }
    } else if (mode == "api") {
        // members the reference callers touch directly
        bloom_filter *index1 = new bloom_filter(64 * MB, 5, 0, 0.01);
        sdbf_set *set1 = new sdbf_set(index1);
        if (set1->index != index1 || index1->bit_mask != 0x1FFFFFFFu || index1->bf_size != 64u * MB)
            return 4;
        uint32_t h[5] = {1, 2, 3, 4, 5};
        if (!index1->insert_sha1(h) || index1->insert_sha1(h) || !index1->query_sha1(h) || index1->elem_count() != 1)
            return 5;
        std::vector<char> buf(4096);
        for (size_t i = 0; i < buf.size(); i++)
            buf[i] = (char)((i * 2654435761u) >> 13);
        sdbf *s = new sdbf("buf", buf.data(), 0, buf.size(), info);
        if (s->big_filters == NULL || s->filter_count() < 1 || s->max_elem != 160)
            return 6;
        sdbf_set *set2 = new sdbf_set();
        set2->add(s);
        set2->vector_init();
        if (set2->filter_count() != s->filter_count() || set2->bf_vector->size() != s->filter_count())
            return 7;
        uint8_t *f = s->clone_filter(0);
        uint16_t h0 = s->hamming[0];
        s->fast();
        uint32_t ones = 0;
        for (int i = 0; i < 64; i++)
            ones += (uint32_t)__builtin_popcount(s->buffer[i]);
        free(f);
        if (s->hamming[0] != ones || ones > h0)
            return 8;
        std::cout << "api ok " << sdbf::get_elem_count(s, 0) << std::endl;
        delete set2;
        delete set1;
        delete index1;
    } else if (mode == "idxw" && argc == 6) {
        // bloom_filter as an index (sdhash_threads.cc:409,430-437): inserts of
        // xorshift64 words, then write_out
        bloom_filter *ix = new bloom_filter((uint64_t)atoll(argv[2]), 5, 0, 0.01);
        uint64_t x = (uint64_t)atoll(argv[4]);
        for (long k = 0; k < atol(argv[3]); k++) {
            uint32_t h[5];
            for (int j = 0; j < 5; j++) {
                x ^= x << 13;
                x ^= x >> 7;
                x ^= x << 17;
                h[j] = (uint32_t)x;
            }
            ix->insert_sha1(h);
        }
        ix->set_name("set.sdbf");
        if (ix->write_out(argv[5]) != 0)
            return 9;
        std::cout << ix->elem_count() << std::endl;
        delete ix;
    } else if (mode == "idxr" && argc == 4) {
        bloom_filter *ix;
        try {
            ix = new bloom_filter(std::string(argv[2]));
        } catch (int e) {
            std::cout << "throw " << e << std::endl;
            return 0;
        }
        std::ofstream os(argv[3], std::ios::binary);
        os.write((const char *)ix->bf, (std::streamsize)ix->bf_size);
        std::cout << ix->bf_size << " " << ix->elem_count() << " " << ix->bit_mask << " " << ix->name() << std::endl;
        delete ix;
    } else {
        return 2;
    }
    free(info);
    return 0;
}
