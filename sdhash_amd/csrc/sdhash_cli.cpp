// sdhash-amd: the sdhash command line (sdhash-src/sdhash.cc) for the hashing
// and comparison paths, running on the gfx950 engine.
//
//   sdhash-amd [options] files...      hash; print digests
//   sdhash-amd -g files...             hash, then compare all pairs
//   sdhash-amd -c a.sdbf [b.sdbf]      compare digests of one or two files
//
// Options: -r/--deep, -g/--gen-compare, -c/--compare, -t/--threshold N,
// -s/--sample-size N, -b/--block-size KB, -p/--threads N (accepted; the
// device does the work),
// -z/--segment-size MB, -o/--output BASE, --separator pipe|csv|tab,
// --hash-name NAME (stdin), --fast, --validate, --index, --index-search DIR,
// --verbose, --version, -h/--help.
// File policy as sdhash.cc:502-618: inputs below 16 MiB are stream-hashed
// unless -b > 0; larger ones use 16 KiB blocks unless -b 0 (or -b N KiB),
// cut into segments named "<file>.NNNNM" (sdhash_threads.cc:300-373).
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "sdbf_amd.h"
#include "hash_files.h"

namespace {

using namespace sdbfx_files;

struct Opts {
    bool deep = false, gen = false, cmp = false, validate = false, verbose = false, fast = false, index = false;
    int32_t threshold = 1;
    uint32_t sample = 0;           // -s: filters of each query digest compared (0 = all)
    int32_t dd = -1;               // KiB; -1 = automatic
    uint64_t segment = 128 * MB;
    std::string output, separator = "pipe", hash_name, index_dir;
    std::vector<std::string> inputs;
};

void usage()
{
    std::cout << "sdhash-amd (sdbf engine for AMD MI355X)\n\n"
                 "Usage: sdhash-amd <options> <files>\n"
                 "  -r [ --deep ]           generate SDBFs from directories and files\n"
                 "  -c [ --compare ]        compare SDBFs in file, or two SDBF files\n"
                 "  -g [ --gen-compare ]    compare all pairs in source data\n"
                 "  -t [ --threshold ] N    only show results >= N (default 1)\n"
                 "  -s [ --sample-size ] N  sample N filters for comparisons (default 0)\n"
                 "  -b [ --block-size ] N   hash input files in N KiB blocks\n"
                 "  -p [ --threads ] N      accepted for compatibility\n"
                 "  -z [ --segment-size ] N segment size in MiB (default 128)\n"
                 "  -o [ --output ] BASE    write BASE.sdbf / BASE.compare\n"
                 "  --separator S           pipe, csv or tab\n"
                 "  --hash-name NAME        name of the digest of stdin\n"
                 "  --fast                  shrink sdbf filters for speedup\n"
                 "  --validate              parse SDBF files and report their contents\n"
                 "  --index                 generate indexes while hashing (needs -o)\n"
                 "  --index-search DIR      search the reference indexes in DIR\n"
                 "  --verbose               progress output\n";
}

bool parse(int argc, char **argv, Opts &o)
{
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i], val;
        auto need = [&](const char *optname) -> bool {
            size_t eq = a.find('=');
            if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
                val = a.substr(eq + 1);
                return true;
            }
            if (a.size() > 2 && a[0] == '-' && a[1] != '-') {
                val = a.substr(2);
                return true;
            }
            if (i + 1 >= argc) {
                std::cerr << "sdhash-amd: missing value for " << optname << std::endl;
                return false;
            }
            val = argv[++i];
            return true;
        };
        auto is = [&](const char *s, const char *l) {
            const std::string L = l;
            return (s && a.rfind(s, 0) == 0 && (a.size() == 2 || (a.size() > 2 && std::string(s).size() == 2 &&
                                                                   strchr("tbpzos", s[1]))) ) ||
                   a == L || a.rfind(L + "=", 0) == 0;
        };
        if (a == "-h" || a == "--help") {
            usage();
            exit(0);
        } else if (a == "--version") {
            std::cout << "sdhash-amd 4.0-compatible (gfx950)" << std::endl;
            exit(0);
        } else if (a == "-r" || a == "--deep") {
            o.deep = true;
        } else if (a == "-g" || a == "--gen-compare") {
            o.gen = true;
        } else if (a == "-c" || a == "--compare") {
            o.cmp = true;
        } else if (a == "--validate") {
            o.validate = true;
        } else if (a == "--index") {
            o.index = true;
        } else if (is(nullptr, "--index-search")) {
            if (!need("index-search"))
                return false;
            o.index_dir = val;
        } else if (a == "--fast") {
            o.fast = true;   // sdhash.cc:268-273
        } else if (a == "--verbose") {
            o.verbose = true;
        } else if (is("-t", "--threshold")) {
            if (!need("threshold"))
                return false;
            o.threshold = atoi(val.c_str());
        } else if (is("-s", "--sample-size")) {
            if (!need("sample-size"))
                return false;
            o.sample = (uint32_t)strtoul(val.c_str(), nullptr, 10);
        } else if (is("-b", "--block-size")) {
            if (!need("block-size"))
                return false;
            o.dd = atoi(val.c_str());
        } else if (is("-p", "--threads")) {
            if (!need("threads"))
                return false;
        } else if (is("-z", "--segment-size")) {
            if (!need("segment-size"))
                return false;
            o.segment = strtoull(val.c_str(), nullptr, 10) * MB;
        } else if (is("-o", "--output")) {
            if (!need("output"))
                return false;
            o.output = val;
        } else if (is(nullptr, "--separator")) {
            if (!need("separator"))
                return false;
            o.separator = val;
        } else if (is(nullptr, "--hash-name")) {
            if (!need("hash-name"))
                return false;
            o.hash_name = val;
        } else if (a.size() > 1 && a[0] == '-' && a != "-") {
            std::cerr << "sdhash-amd: unknown option " << a << std::endl;
            return false;
        } else {
            o.inputs.push_back(a);
        }
    }
    // sdhash.cc:209-218
    if (o.dd > 16)
        o.dd = 16;
    if (o.dd < -1)
        o.dd = -1;
    return true;
}

char sep_char(const std::string &s)
{
    if (s == "csv")
        return ',';
    if (s == "tab" || s == "tabs")
        return '\t';
    return '|';
}

bool regular(const std::string &p, uint64_t *size)
{
    struct stat st;
    if (stat(p.c_str(), &st) != 0 || !S_ISREG(st.st_mode))
        return false;
    *size = (uint64_t)st.st_size;
    return true;
}

void walk(const std::string &dir, std::vector<std::string> &out)
{
    DIR *d = opendir(dir.c_str());
    if (!d)
        return;
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) {
        if (!strcmp(e->d_name, ".") || !strcmp(e->d_name, ".."))
            continue;
        names.push_back(dir + "/" + e->d_name);
    }
    closedir(d);
    for (const auto &n : names) {
        struct stat st;
        if (lstat(n.c_str(), &st) != 0)
            continue;
        if (S_ISDIR(st.st_mode))
            walk(n, out);
        else if (S_ISREG(st.st_mode))
            out.push_back(n);
    }
}

int write_file(const std::string &s, const std::string &fn)
{
    std::ofstream os(fn, std::ios::binary);
    if (!os) {
        std::cerr << "sdhash: ERROR cannot write to file " << fn << std::endl;
        return -1;
    }
    os << s;
    return 0;
}

// sdhash_threads.cc:376-492 (hash_index_stringlist), run in the order given
// (sdhash -p 1): files are grouped, a group closing once the sum of its file
// sizes passes 640 MB or at the last file; each group is hashed against a new
// 64 MiB index into <output><n>.sdbf, whose index goes to <output><n>.sdbf.idx.
// Files below 16 MiB are stream-hashed; larger ones use 16 KiB blocks and
// 128 MiB segments.  As in the reference, every group after the first also
// starts with the first file of the first group (its file list is refilled
// from the second slot on), and the numbering restarts at 0 for the large
// files, whose outputs replace the small files' of the same number.
int hash_index_list(const std::vector<std::string> &files, const std::string &output)
{
    std::vector<std::string> small, large;
    std::vector<uint64_t> ssz, lsz;
    for (const auto &f : files) {
        uint64_t sz;
        if (!regular(f, &sz))
            continue;
        if (sz < 16 * MB)
            small.push_back(f), ssz.push_back(sz);
        else
            large.push_back(f), lsz.push_back(sz);
    }
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<std::string> &lst = pass ? large : small;
        const std::vector<uint64_t> &szs = pass ? lsz : ssz;
        std::vector<std::string> group;
        uint64_t total = 0;
        int count = 0;
        for (size_t i = 0; i < lst.size(); i++) {
            group.push_back(lst[i]);
            total += szs[i];
            if (!(total > 640 * MB || i == lst.size() - 1))
                continue;
            bloom_filter *index1 = new bloom_filter(64 * MB, 5, 0, 0.01);
            index_info info{index1, nullptr, nullptr, false, false, false};
            sdbf_set *set1 = new sdbf_set(index1);
            std::vector<Job> jobs;
            for (const auto &f : group) {
                std::vector<uint8_t> d;
                if (!read_file(f, d))
                    continue;
                if (pass) {
                    // large files one at a time (bounded host memory); the
                    // index carries over from call to call
                    segment_jobs(f, d, 128 * MB, jobs);
                    run_jobs(jobs, 16 * KB, set1, &info);
                } else {
                    const uint64_t got = d.size();
                    jobs.push_back(Job{f, std::move(d), got});
                }
            }
            if (!jobs.empty())
                run_jobs(jobs, 0, set1, &info);
            const std::string nm = output + std::to_string(count) + ".sdbf";
            count++;
            if (write_file(set1->to_string(), nm) != 0)
                return -1;
            set1->index->set_name(nm);
            if (set1->index->write_out(nm + ".idx") == -2) {
                std::cerr << "sdhash: ERROR cannot write to file " << nm + ".idx" << std::endl;
                return -1;
            }
            for (uint64_t k = 0; k < set1->size(); k++)
                delete set1->at((uint32_t)k);
            delete set1;
            delete index1;
            group.assign(1, lst[0]);
            total = 0;
        }
    }
    return 0;
}

// --index-search DIR (sdhash.cc:241-264): every <name>.idx file in DIR (in
// name order; the reference takes the directory's own order) is loaded with
// the set of DIR/<name> that it indexes, named by that path.
bool load_index_dir(const std::string &dir, std::vector<sdbf_set *> &setlist)
{
    DIR *d = opendir(dir.c_str());
    if (!d)
        return true;
    std::vector<std::string> names;
    while (dirent *e = readdir(d)) {
        const std::string n = e->d_name;
        uint64_t sz;
        if (n.size() > 4 && n.compare(n.size() - 4, 4, ".idx") == 0 && regular(dir + "/" + n, &sz))
            names.push_back(n);
    }
    closedir(d);
    std::sort(names.begin(), names.end());
    for (const auto &n : names) {
        bloom_filter *b = new bloom_filter(dir + "/" + n);
        sdbf_set *s = new sdbf_set((dir + "/" + n.substr(0, n.size() - 4)).c_str());
        s->index = b;
        setlist.push_back(s);
    }
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    Opts o;
    if (!parse(argc, argv, o))
        return 1;
    sdbf::config->warnings = o.verbose ? 1 : 0;
    const char sep = sep_char(o.separator);
    // sdhash.cc:195-198
    if (o.index && o.output.empty()) {
        std::cerr << "sdhash:  ERROR: indexing requires output base filename " << std::endl;
        return -1;
    }
    try {
        std::vector<sdbf_set *> setlist;
        index_info search_info{nullptr, nullptr, &setlist, false, false, false};
        index_info *info = nullptr;
        if (!o.index_dir.empty()) {
            // sdhash.cc:241-250
            if (o.dd == 0) {
                std::cerr << "ERROR: Index searching only supported in block mode" << std::endl;
                return 0;
            }
            if (o.dd == -1)
                o.dd = 16;
            load_index_dir(o.index_dir, setlist);
            info = &search_info;
        }
        if (o.cmp) {
            if (o.inputs.size() == 1) {
                sdbf_set *s1 = new sdbf_set(o.inputs[0].c_str());
                s1->set_separator(sep);
                if (!o.output.empty())
                    write_file(s1->compare_all_quiet(o.threshold, 0, o.fast), o.output + ".compare");
                else
                    s1->compare_all(o.threshold, o.fast);
            } else if (o.inputs.size() == 2) {
                sdbf_set *s1 = new sdbf_set(o.inputs[0].c_str());
                sdbf_set *s2 = new sdbf_set(o.inputs[1].c_str());
                s1->set_separator(sep);
                if (!o.output.empty())
                    write_file(s1->compare_to_quiet(s2, o.threshold, o.sample, 0, o.fast), o.output + ".compare");
                else
                    s1->compare_to(s2, o.threshold, o.sample, o.fast);
            } else {
                std::cerr << "sdhash: ERROR: Comparison requires 1 or 2 arguments." << std::endl;
                return -1;
            }
            return 0;
        }
        if (o.validate) {
            for (const auto &f : o.inputs) {
                uint64_t sz;
                if (!regular(f, &sz)) {
                    std::cout << "sdhash: ERROR file " << f << " not readable or not found." << std::endl;
                    continue;
                }
                sdbf_set s(f.c_str());
                std::cout << "sdhash: file " << f << " SDBFs valid, contains " << s.size() << " hashes." << std::endl;
                std::cout << "contains " << s.filter_count() << " bloom filters." << std::endl;
                std::cout << s.input_size() << " total data size. " << std::endl;
            }
            return 0;
        }
        if (o.index) {
            // sdhash.cc:556-561: files below 16 MiB, then the larger ones
            std::vector<std::string> files, small, large;
            for (const auto &in : o.inputs) {
                std::vector<std::string> found;
                struct stat st;
                if (o.deep && stat(in.c_str(), &st) == 0 && S_ISDIR(st.st_mode))
                    walk(in, found);
                uint64_t sz;
                if (regular(in, &sz))
                    found.push_back(in);
                for (const auto &f : found)
                    if (regular(f, &sz))
                        (sz < 16 * MB ? small : large).push_back(f);
            }
            const int a = hash_index_list(small, o.output), b = hash_index_list(large, o.output);
            return a + b;
        }
        const bool keep = o.gen || !o.output.empty() || info;
        sdbf_set *set1 = keep ? new sdbf_set() : nullptr;
        std::vector<Job> stream_jobs, block_jobs;
        uint32_t block_bytes = 0;
        if (o.inputs.size() == 1 && o.inputs[0] == "-") {
            // stdin, always block mode and always in segments (sdhash.cc:
            // 460-472: a segment size of 0 becomes 128 MiB; sdhash_threads.cc:
            // 189-233)
            if (o.segment == 0)
                o.segment = 128 * MB;
            const uint32_t dd = o.dd > 0 ? (uint32_t)o.dd : 16;
            const std::string base = o.hash_name.empty() ? "stdin" : o.hash_name;
            std::vector<uint8_t> buf(o.segment);
            for (uint64_t j = 0;; j++) {
                std::fill(buf.begin(), buf.end(), 0);
                const size_t got = fread(buf.data(), 1, buf.size(), stdin);
                std::stringstream ns;
                ns.fill('0');
                ns << base << "." << std::setw(4) << j * o.segment / MB << "M";
                std::vector<Job> jobs{Job{ns.str(), std::vector<uint8_t>(buf.begin(), buf.begin() + got), got}};
                run_jobs(jobs, dd * KB, set1, info);
                if (got != buf.size())
                    break;
            }
        } else {
            std::vector<std::string> small, large;
            for (const auto &in : o.inputs) {
                std::vector<std::string> found;
                uint64_t sz;
                if (o.deep) {
                    struct stat st;
                    if (stat(in.c_str(), &st) == 0 && S_ISDIR(st.st_mode))
                        walk(in, found);
                }
                if (regular(in, &sz))
                    found.push_back(in);
                for (const auto &f : found) {
                    if (!regular(f, &sz))
                        continue;
                    (sz < 16 * MB ? small : large).push_back(f);
                }
            }
            // sdhash.cc:554-618
            auto add_stream = [&](const std::string &f) {
                std::vector<uint8_t> d;
                if (read_file(f, d))
                    stream_jobs.push_back(Job{f, std::move(d), 0}), stream_jobs.back().got =
                                                                        stream_jobs.back().data.size();
            };
            auto add_block = [&](const std::string &f, uint32_t kb) {
                std::vector<uint8_t> d;
                if (read_file(f, d))
                    segment_jobs(f, d, o.segment, block_jobs);
                block_bytes = kb * KB;
            };
            for (const auto &f : small) {
                if (o.dd < 1)
                    add_stream(f);
                else
                    add_block(f, (uint32_t)o.dd);
            }
            if (!stream_jobs.empty())
                run_jobs(stream_jobs, 0, set1, info);
            if (!block_jobs.empty())
                run_jobs(block_jobs, block_bytes, set1, info);
            for (const auto &f : large) {
                if (o.dd == 0)
                    add_stream(f);
                else
                    add_block(f, o.dd == -1 ? 16 : (uint32_t)o.dd);
                // large inputs are digested one file at a time (bounded host memory)
                if (!stream_jobs.empty())
                    run_jobs(stream_jobs, 0, set1, info);
                if (!block_jobs.empty())
                    run_jobs(block_jobs, block_bytes, set1, info);
            }
        }
        if (o.gen) {
            set1->set_separator(sep);
            if (!o.output.empty())
                write_file(set1->compare_all_quiet(o.threshold, 0, o.fast), o.output + ".compare");
            else
                set1->compare_all(o.threshold, o.fast);
        } else if (info) {
            // sdhash.cc:640-652
            if (!o.output.empty())
                write_file(set1->index_results(), o.output + ".idx-result");
            else
                std::cerr << set1->index_results();
        } else if (!o.output.empty()) {
            write_file(set1->to_string(), o.output + ".sdbf");
        }
    } catch (int e) {
        if (e == -2) {
            std::cerr << "sdhash: ERROR: invalid SDBF input" << std::endl;
            return -2;
        }
        std::cerr << "sdhash: ERROR (" << e << ")" << std::endl;
        return -1;
    }
    return 0;
}
