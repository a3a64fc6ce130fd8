// File hashing shared by the command line (sdhash_cli.cpp) and the server
// handler (sdhash_srv.cpp): reading files, the segmentation of
// sdhash-src/sdhash_threads.cc:300-373 and one batched device call per list
// of digests.  Header-only, built into each.
#ifndef SDBFX_HASH_FILES_H
#define SDBFX_HASH_FILES_H
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "sdbf_amd.h"

namespace sdbfx_files {

const uint64_t KB = 1024, MB = 1024 * 1024;

inline bool read_file(const std::string &p, std::vector<uint8_t> &data)
{
    std::ifstream is(p, std::ios::binary);
    if (!is)
        return false;
    is.seekg(0, std::ios::end);
    const std::streamoff n = is.tellg();
    is.seekg(0);
    data.resize((size_t)n);
    is.read((char *)data.data(), n);
    return is.gcount() == n;
}

// One digest to compute: name, bytes (zero-padded to `len`), actual size read.
struct Job {
    std::string name;
    std::vector<uint8_t> data;
    uint64_t got;
};

// Segmentation of sdhash_threads.cc:300-373 for one file of `filesize`
// bytes, with its byte-exact quirk: the last segment after a short tail is
// read with the tail counted twice (zero-filled past the end of the file).
inline void segment_jobs(const std::string &name, const std::vector<uint8_t> &file, uint64_t chunk,
                  std::vector<Job> &jobs)
{
    const uint64_t filesize = file.size();
    if (!(filesize > chunk && chunk > 0)) {
        jobs.push_back(Job{name, file, filesize});
        return;
    }
    int64_t chunks = (int64_t)(filesize / chunk);
    int tailflag = 0;
    if (filesize - chunk <= 512)
        chunks = 0;
    if (filesize - (uint64_t)chunks * chunk <= 512) {
        chunks--;
        tailflag = 1;
    }
    uint64_t pos = 0;
    for (int64_t j = 0; j <= chunks; j++) {
        std::stringstream ns;
        ns << name;
        if (j > 0 || chunks > 0) {
            ns.fill('0');
            ns << "." << std::setw(4) << (uint64_t)j * chunk / MB << "M";
        }
        uint64_t csize;
        if (j == chunks) {
            csize = filesize - (uint64_t)chunks * chunk;
            if (tailflag)
                csize = csize + (filesize - (uint64_t)(chunks + 1) * chunk);
        } else {
            csize = chunk;
        }
        const uint64_t got = std::min<uint64_t>(csize, filesize - pos);
        Job jb{ns.str(), std::vector<uint8_t>(csize, 0), got};
        memcpy(jb.data.data(), file.data() + pos, got);
        pos += got;
        jobs.push_back(std::move(jb));
    }
}

// Digests jobs in one device batch and appends them to `set` (or prints);
// with info, against its indexes, in order.
inline int run_jobs(std::vector<Job> &jobs, uint32_t dd_bytes, sdbf_set *set, index_info *info = nullptr)
{
    std::vector<std::string> names;
    std::vector<const uint8_t *> bufs;
    std::vector<uint64_t> lens, got;
    for (size_t i = 0; i < jobs.size(); i++) {
        if (jobs[i].got < 512)   // the constructors' -3: skipped silently
            continue;
        names.push_back(jobs[i].name);
        bufs.push_back(jobs[i].data.data());
        lens.push_back(jobs[i].data.size());
        // the stream constructor's orig_file_size is the bytes it read
        // (gcount, sdbf_class.cc:77-85), not the zero-padded length
        got.push_back(jobs[i].got);
    }
    std::vector<sdbf *> out;
    if (!bufs.empty())
        sdbf::hash_batch(names, bufs, lens, dd_bytes, out, &got, info);
    for (size_t k = 0; k < out.size(); k++) {
        if (!out[k])
            continue;
        if (set)
            set->add(out[k]);
        else {
            std::cout << out[k];
            delete out[k];
        }
    }
    jobs.clear();
    return 0;
}

}  // namespace sdbfx_files

#endif
