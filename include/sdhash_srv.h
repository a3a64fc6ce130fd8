// sdhash_srv.h -- the request handler of sdhash-server (sdhashsrvHandler,
// sdhash-server/src/sdhash-srv-main.cc:52-660) on the gfx950 engine, with the
// methods and signatures of the Thrift service interface sdhashsrvIf
// (sdhash-server/gen-cpp/sdhashsrv.h:15-39, from sdhashsrv.thrift).  Thrift
// itself is not part of this build: a server binds this class by forwarding
// each generated sdhashsrvIf method to the method of the same name.
//
// The loaded sets stay resident in device memory (sdbf_set keeps a device
// copy once compared), so every compareAll / compareTwo on them runs the
// comparison kernels without re-uploading, and hashString digests its whole
// file list in batched device calls rather than file by file.
//
// Differences from the reference: set and result IDs are handed out in
// sequence from 1 (the reference draws them from a time-seeded mt19937 in
// [1, 131070]); shutdown() only sets shutting_down() (there is no transport
// to stop); the work of a oneway call is done before it returns.
#ifndef SDHASH_SRV_H
#define SDHASH_SRV_H

#include <stdint.h>
#include <time.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sdbf_amd.h"

class sdhashsrvHandler {
public:
    // loads every readable .sdbf set in `home` (with its .idx index when one
    // exists) and registers it under its file stem (sdhash-srv-main.cc:59-104)
    sdhashsrvHandler(uint32_t thread_count, std::string *home, std::string *sources);
    ~sdhashsrvHandler();

    void ping();
    void setList(std::string &_return, const bool json);
    void displaySet(std::string &_return, const int32_t num1);
    void displayContents(std::string &_return, const int32_t num1);
    void compareAll(const int32_t num1, const int32_t threshold, const int32_t resultID);
    void compareTwo(const int32_t num1, const int32_t num2, const int32_t threshold, const int32_t sample,
                    const int32_t resultID);
    int32_t loadSet(const std::string &filename, const int32_t hashsetID);
    int32_t saveSet(const int32_t num1, const std::string &filename);
    int32_t createHashsetID();
    void hashString(const std::string &setname, const std::vector<std::string> &filenames, const int32_t blocksize,
                    const int32_t hashsetID, const int32_t searchIndex);
    void displaySourceList(std::vector<std::string> &_return);
    void getHashsetName(std::string &_return, const int32_t num1);
    void displayResult(std::string &_return, const int32_t resultID);
    void displayResultsList(std::string &_return, const std::string &user, const bool json);
    int32_t createResultID(const std::string &user);
    void displayResultStatus(std::string &_return, const int32_t resultID);
    void displayResultDuration(std::string &_return, const int32_t resultID);
    void displayResultInfo(std::string &_return, const int32_t resultID);
    bool removeResult(const int32_t resultID);
    bool saveResult(const int32_t resultID, const std::string &result, const std::string &info);
    void shutdown();

    bool shutting_down() const { return down; }

private:
    int32_t next_id();
    int32_t add_set(sdbf_set *set, const std::string &name, int32_t setID);
    sdbf_set *get_set(int32_t setID);
    std::string get_set_name(int32_t setID);
    void add_request(int32_t resultID, const std::string &req);
    void add_result(int32_t resultID, const std::string &res);
    std::string get_result(int32_t resultID);
    std::string get_request(int32_t resultID);
    sdbf_set *hash_stringlist(const std::vector<std::string> &filenames, uint32_t dd_block_size, index_info *info);
    int32_t hash_index_stringlist(const std::vector<std::string> &filenames, const std::string &output_name);

    std::string source_directory, home_directory;
    uint32_t processing_thread_count;
    std::multimap<std::string, int32_t> users;
    std::map<int32_t, std::string> requests, results;
    std::map<int32_t, time_t> starttime, endtime;
    std::map<int32_t, sdbf_set *> setcollection;
    std::vector<sdbf_set *> setlist;        // the sets that have an index
    std::vector<bloom_filter *> indexlist;
    int32_t last_id = 0;
    bool down = false;
    std::recursive_mutex mu;
};

#endif
