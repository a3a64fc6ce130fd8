// sdhashsrvHandler on the gfx950 engine: the request handler of
// sdhash-server (sdhash-server/src/sdhash-srv-main.cc:52-660, with the file
// hashing of set_list.cc:108-152), without the Thrift transport.
#include "sdhash_srv.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>

#include "hash_files.h"

using namespace sdbfx_files;

namespace {

const uint64_t GB = 1024ull * MB;

bool is_regular(const std::string &p, uint64_t *size = nullptr)
{
    struct stat st;
    if (stat(p.c_str(), &st) != 0 || !S_ISREG(st.st_mode))
        return false;
    if (size)
        *size = (uint64_t)st.st_size;
    return true;
}

bool is_dir(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// entries of a directory in name order (the reference takes the
// filesystem's own order); recursive lists regular files below it
void list_dir(const std::string &dir, bool recursive, std::vector<std::string> &out)
{
    DIR *d = opendir(dir.c_str());
    if (!d)
        return;
    std::vector<std::string> names;
    while (dirent *e = readdir(d))
        if (strcmp(e->d_name, ".") && strcmp(e->d_name, ".."))
            names.push_back(e->d_name);
    closedir(d);
    std::sort(names.begin(), names.end());
    for (const auto &n : names) {
        const std::string p = dir + "/" + n;
        if (is_regular(p))
            out.push_back(p);
        else if (recursive && is_dir(p))
            list_dir(p, true, out);
    }
}

std::string stem(const std::string &path)
{
    std::string f = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
    const size_t dot = f.find_last_of('.');
    return dot == std::string::npos || dot == 0 ? f : f.substr(0, dot);
}

// sizes as the listings print them (sdhash-srv-main.cc:131-141)
std::string human(uint64_t s)
{
    std::stringstream o;
    if (s > GB)
        o << s / GB << 'G';
    else if (s > MB)
        o << s / MB << 'M';
    else
        o << s / KB << 'K';
    return o.str();
}

}  // namespace

sdhashsrvHandler::sdhashsrvHandler(uint32_t thread_count, std::string *home, std::string *sources)
    : source_directory(sources ? *sources : std::string()), home_directory(home ? *home : std::string()),
      processing_thread_count(thread_count)
{
    uint32_t j = 0;
    std::vector<std::string> files;
    list_dir(home_directory, false, files);
    for (const auto &p : files) {
        try {
            std::cerr << "-> Loading file " << p << " ..... ";
            sdbf_set *tmp = new sdbf_set(p.c_str());
            if (!tmp->empty()) {
                j++;
                std::cerr << "OK" << std::endl;
                if (is_regular(p + ".idx")) {
                    bloom_filter *ix = new bloom_filter(p + ".idx");
                    indexlist.push_back(ix);
                    tmp->index = ix;
                    std::cerr << "-> Loaded index " << p << "..... OK" << std::endl;
                }
                add_set(tmp, stem(p), next_id());
            } else {
                std::cerr << "File empty" << std::endl;
                delete tmp;
            }
        } catch (int e) {
            std::cerr << (e == -2 ? "File format invalid" : e == -3 ? "File too small" : "File not accessible")
                      << std::endl;
        }
    }
    std::cerr << "sdhash-srv loaded " << j << " sets." << std::endl;
}

sdhashsrvHandler::~sdhashsrvHandler()
{
    for (auto &kv : setcollection) {
        for (uint64_t k = 0; k < kv.second->size(); k++)
            delete kv.second->at((uint32_t)k);
        delete kv.second;
    }
    for (bloom_filter *b : indexlist)
        delete b;
}

int32_t sdhashsrvHandler::next_id()
{
    std::lock_guard<std::recursive_mutex> g(mu);
    return ++last_id;
}

void sdhashsrvHandler::ping() { std::cout << "ping" << std::endl; }

void sdhashsrvHandler::setList(std::string &_return, const bool json)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    std::stringstream out;
    if (json)
        out << "{ \"aaData\": [";
    int i = 0;
    for (auto it = setcollection.begin(); it != setcollection.end(); ++it, ++i) {
        const std::string sz = human(it->second->input_size());
        if (json) {
            if (i)
                out << ",";
            out << "[\"" << it->first << "\",\"" << it->second->name() << "\",\"" << sz << "\",\""
                << it->second->size() << "\"]";
        } else {
            out << it->first << " " << it->second->name() << " " << sz << "\n";
        }
    }
    if (json)
        out << "] }";
    if (setcollection.empty())
        _return = json ? "[\"sdhash-srv ready, no sets loaded\"]" : "sdhash-srv ready, no sets loaded\n";
    else
        _return = out.str();
}

void sdhashsrvHandler::displaySet(std::string &_return, const int32_t num1)
{
    sdbf_set *t = get_set(num1);
    if (!t) {
        _return = "Error: not found\n";
        return;
    }
    std::stringstream out;
    for (uint64_t i = 0; i < t->size(); i++)
        out << t->at((uint32_t)i)->name() << " " << human(t->at((uint32_t)i)->input_size()) << "\n";
    _return = out.str();
}

void sdhashsrvHandler::displayContents(std::string &_return, const int32_t num1)
{
    sdbf_set *t = get_set(num1);
    _return = t ? t->to_string() : "Error: not found\n";
}

void sdhashsrvHandler::compareAll(const int32_t num1, const int32_t threshold, const int32_t resultID)
{
    sdbf_set *s = get_set(num1);
    if (!s)
        return;
    add_request(resultID, get_set_name(num1) + " all hashes");
    add_result(resultID, s->compare_all_quiet(threshold, (int32_t)processing_thread_count, false));
}

void sdhashsrvHandler::compareTwo(const int32_t num1, const int32_t num2, const int32_t threshold,
                                  const int32_t sample, const int32_t resultID)
{
    sdbf_set *a = get_set(num1), *b = get_set(num2);
    if (!a || !b)
        return;
    add_request(resultID, get_set_name(num1) + " -- " + get_set_name(num2));
    add_result(resultID, a->compare_to_quiet(b, threshold, (uint32_t)sample, (int32_t)processing_thread_count, false));
}

// sdhash-srv-main.cc:267-304: the set gets a new ID (hashsetID is not used)
int32_t sdhashsrvHandler::loadSet(const std::string &filename, const int32_t)
{
    try {
        sdbf_set *tmp = new sdbf_set(filename.c_str());
        if (is_regular(filename + ".idx")) {
            bloom_filter *ix = new bloom_filter(filename + ".idx");
            std::lock_guard<std::recursive_mutex> g(mu);
            indexlist.push_back(ix);
            tmp->index = ix;
        }
        if (tmp->empty()) {
            std::cerr << "File not found or empty: " << filename << std::endl;
            delete tmp;
            return -1;
        }
        return add_set(tmp, filename, next_id());
    } catch (int e) {
        std::cerr << (e == -2 ? "File format invalid: " : e == -3 ? "File too short: " : "File not accessible: ")
                  << filename << std::endl;
        return -1;
    }
}

int32_t sdhashsrvHandler::saveSet(const int32_t num1, const std::string &filename)
{
    sdbf_set *s = get_set(num1);
    if (!s)
        return -1;
    std::ofstream os(filename.c_str(), std::ios::out);
    if (!os.is_open())
        return -1;
    os << s;
    return 0;
}

int32_t sdhashsrvHandler::createHashsetID() { return next_id(); }

// set_list.cc:108-152: files below 16 MiB stream-hashed (or in dd_block_size
// KiB blocks), larger ones in 16 KiB blocks (or dd_block_size), both in
// 128 MiB segments; each list goes to the device as one batch
sdbf_set *sdhashsrvHandler::hash_stringlist(const std::vector<std::string> &filenames, uint32_t dd, index_info *info)
{
    sdbf_set *set1 = new sdbf_set(info ? info->index : nullptr);
    std::vector<std::string> small, large;
    for (const auto &f : filenames) {
        uint64_t sz;
        if (is_regular(f, &sz))
            (sz < 16 * MB ? small : large).push_back(f);
    }
    std::vector<Job> jobs;
    for (const auto &f : small) {
        std::vector<uint8_t> d;
        if (!read_file(f, d))
            continue;
        if (dd < 1) {
            const uint64_t got = d.size();
            jobs.push_back(Job{f, std::move(d), got});
        } else {
            segment_jobs(f, d, 128 * MB, jobs);
        }
    }
    if (!jobs.empty())
        run_jobs(jobs, dd < 1 ? 0 : dd * (uint32_t)KB, set1, info);
    for (const auto &f : large) {
        std::vector<uint8_t> d;
        if (!read_file(f, d))
            continue;
        segment_jobs(f, d, 128 * MB, jobs);
        run_jobs(jobs, (dd < 1 ? 16 : dd) * (uint32_t)KB, set1, info);
    }
    return set1;
}

// sdhash-srv-main.cc:667-787: as sdhash --index, with the server's own
// quirks kept: a group closes once its file sizes pass 640 MB or at the last
// file, and is hashed without that closing file (the file count passed is
// one short) and, after the first group, with the first group's first file
// in front (the list is refilled from its second slot); small-file outputs
// go to the home directory, large-file ones to output_name as given; every
// group's set is registered with its index, which later searches use.
int32_t sdhashsrvHandler::hash_index_stringlist(const std::vector<std::string> &filenames,
                                                const std::string &output_name)
{
    std::vector<std::string> small, large;
    std::vector<uint64_t> ssz, lsz;
    for (const auto &f : filenames) {
        uint64_t sz;
        if (!is_regular(f, &sz))
            continue;
        if (sz < 16 * MB)
            small.push_back(f), ssz.push_back(sz);
        else
            large.push_back(f), lsz.push_back(sz);
    }
    int count = 0;
    for (int pass = 0; pass < 2; pass++) {
        const std::vector<std::string> &lst = pass ? large : small;
        const std::vector<uint64_t> &szs = pass ? lsz : ssz;
        std::vector<std::string> slots;   // the reference's file list
        size_t filect = 0;
        uint64_t total = 0;
        for (size_t i = 0; i < lst.size(); filect++, i++) {
            if (slots.size() <= filect)
                slots.resize(filect + 1);
            slots[filect] = lst[i];
            total += szs[i];
            if (!(total > 640 * MB || i == lst.size() - 1))
                continue;
            bloom_filter *index1 = new bloom_filter(64 * MB, 5, 0, 0.01);
            index_info info{index1, nullptr, nullptr, false, false, false};
            std::vector<std::string> group(slots.begin(), slots.begin() + filect);
            sdbf_set *set1;
            if (pass) {
                set1 = new sdbf_set(index1);
                std::vector<Job> jobs;
                for (const auto &f : group) {
                    std::vector<uint8_t> d;
                    if (read_file(f, d)) {
                        segment_jobs(f, d, 128 * MB, jobs);
                        run_jobs(jobs, 16 * (uint32_t)KB, set1, &info);
                    }
                }
            } else {
                set1 = hash_stringlist(group, 0, &info);
            }
            const std::string nm = (pass ? output_name : home_directory + "/" + output_name) + std::to_string(count) +
                                   ".sdbf";
            total = 0;
            filect = 0;
            count++;
            std::ofstream os(nm.c_str(), std::ios::out | std::ios::binary);
            if (!os.is_open()) {
                std::cerr << "sdhash: ERROR cannot write to file " << output_name << std::endl;
                return -1;
            }
            os << set1;
            os.close();
            set1->index->set_name(nm);
            if (set1->index->write_out(nm + ".idx") == -2) {
                std::cerr << "sdhash: ERROR cannot write to file " << nm + ".idx" << std::endl;
                return -1;
            }
            {
                std::lock_guard<std::recursive_mutex> g(mu);
                indexlist.push_back(index1);
            }
            add_set(set1, nm, next_id());
        }
    }
    return 0;
}

// sdhash-srv-main.cc:332-402
void sdhashsrvHandler::hashString(const std::string &setname, const std::vector<std::string> &filenames,
                                  const int32_t blocksize, const int32_t hashsetID, const int32_t searchIndex)
{
    const uint32_t dd = blocksize > 0 ? (uint32_t)std::min(blocksize, 16) : 0;
    if (searchIndex > 0) {
        // 1: search top, 2: search all, 3: search the set; only 3 reports
        // (print_indexes runs when neither flag is set, sdbf_core.cc:184)
        std::vector<sdbf_set *> sl;
        {
            std::lock_guard<std::recursive_mutex> g(mu);
            sl = setlist;
        }
        index_info info{nullptr, &indexlist, &sl, searchIndex != 3, searchIndex != 3, false};
        add_request(hashsetID, setname + " indexes");
        sdbf_set *tmp = hash_stringlist(filenames, dd, &info);
        add_result(hashsetID, tmp->index_results());
        for (uint64_t k = 0; k < tmp->size(); k++)
            delete tmp->at((uint32_t)k);
        delete tmp;
    } else if (searchIndex == 0) {
        sdbf_set *tmp = hash_stringlist(filenames, dd, nullptr);
        if (!tmp->empty()) {
            const std::string savename = home_directory + "/" + setname + ".sdbf";
            add_set(tmp, setname, hashsetID);
            std::ofstream os(savename.c_str(), std::ios::out);
            if (os.is_open())
                os << tmp;
        } else {
            delete tmp;
        }
    } else {
        hash_index_stringlist(filenames, setname);
    }
}

void sdhashsrvHandler::displaySourceList(std::vector<std::string> &_return)
{
    if (is_dir(source_directory))
        list_dir(source_directory, true, _return);
    else
        _return.push_back(source_directory + " is not accessible");
}

void sdhashsrvHandler::getHashsetName(std::string &_return, const int32_t num1) { _return = get_set_name(num1); }

void sdhashsrvHandler::displayResult(std::string &_return, const int32_t resultID) { _return = get_result(resultID); }

void sdhashsrvHandler::displayResultsList(std::string &_return, const std::string &user, const bool json)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    std::stringstream o;
    if (json)
        o << "[";
    auto range = users.equal_range(user);
    for (auto it = range.first; it != range.second; ++it) {
        if (json) {
            if (it != range.first)
                o << ",";
            o << "{\"" << it->second << "\":\"" << get_request(it->second) << "\"}";
        } else {
            o << it->second << " " << get_request(it->second) << std::endl;
        }
    }
    if (json)
        o << "]";
    _return = o.str();
}

int32_t sdhashsrvHandler::createResultID(const std::string &user)
{
    const int32_t id = next_id();
    std::lock_guard<std::recursive_mutex> g(mu);
    users.insert(std::make_pair(user, id));
    return id;
}

// sdhash-srv-main.cc:602-625: "complete (<lines>)", "processing" or "not found"
void sdhashsrvHandler::displayResultStatus(std::string &_return, const int32_t resultID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    const bool s = starttime.count(resultID) != 0, e = endtime.count(resultID) != 0;
    if (s && e)
        _return = "complete (" + std::to_string(std::count(results[resultID].begin(), results[resultID].end(), '\n')) +
                  ")";
    else if (s)
        _return = "processing";
    else
        _return = "not found";
}

void sdhashsrvHandler::displayResultDuration(std::string &_return, const int32_t resultID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    _return = "";
    if (!starttime.count(resultID))
        return;
    const time_t end = endtime.count(resultID) ? endtime[resultID] : time(0);
    _return = std::to_string((long long)(end - starttime[resultID])) + "s";
}

void sdhashsrvHandler::displayResultInfo(std::string &_return, const int32_t resultID)
{
    _return = get_request(resultID);
}

bool sdhashsrvHandler::removeResult(const int32_t resultID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    requests.erase(resultID);
    results.erase(resultID);
    starttime.erase(resultID);
    endtime.erase(resultID);
    for (auto it = users.begin(); it != users.end();)
        it = it->second == resultID ? users.erase(it) : std::next(it);
    return true;
}

// the reference does not keep results given by clients (sdhash-srv-main.cc:490-494)
bool sdhashsrvHandler::saveResult(const int32_t, const std::string &, const std::string &) { return false; }

void sdhashsrvHandler::shutdown()
{
    std::cout << "sdhash-srv shutdown requested by client. Shutting down." << std::endl;
    down = true;
}

int32_t sdhashsrvHandler::add_set(sdbf_set *set, const std::string &name, int32_t setID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    setcollection[setID] = set;
    if (set->index)
        setlist.push_back(set);
    set->set_name(name);
    set->vector_init();
    return setID;
}

sdbf_set *sdhashsrvHandler::get_set(int32_t setID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    auto it = setcollection.find(setID);
    return it == setcollection.end() ? nullptr : it->second;
}

std::string sdhashsrvHandler::get_set_name(int32_t setID)
{
    sdbf_set *s = get_set(setID);
    return s ? s->name() : std::string();
}

void sdhashsrvHandler::add_request(int32_t resultID, const std::string &req)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    requests.insert(std::make_pair(resultID, req));
    starttime.insert(std::make_pair(resultID, time(0)));
}

void sdhashsrvHandler::add_result(int32_t resultID, const std::string &res)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    results.insert(std::make_pair(resultID, res));
    endtime.insert(std::make_pair(resultID, time(0)));
}

std::string sdhashsrvHandler::get_result(int32_t resultID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    if (results.count(resultID) && starttime.count(resultID) && endtime.count(resultID))
        return results[resultID];
    return "";
}

std::string sdhashsrvHandler::get_request(int32_t resultID)
{
    std::lock_guard<std::recursive_mutex> g(mu);
    auto it = requests.find(resultID);
    return it == requests.end() ? std::string() : it->second;
}
