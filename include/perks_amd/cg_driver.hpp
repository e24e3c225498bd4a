// cg_driver.hpp -- the reference's CG driver entry point over the C ABI.
//
// The reference's main (conjugateGradient/main.cu:36-84) parses its command
// line into a CommandLineArgs and calls
//   template <typename ValueT, typename OffsetT, bool baseline, bool cacheMatrix, bool cacheVector>
//   void myTest(int devID, cudaDeviceProp& deviceProp, CommandLineArgs& args);
// (conjugateGradient/cg_driver.cuh:4-5).  A driver written against it includes
// this header and links libperks_amd.so; the call keeps its template
// arguments, command-line flags (--mtx=, --vdata, --iters=, --staticiter,
// --warmup, --check) and report.  The device properties type is the neutral
// perks_device_caps (perks_query_device fills it).  The cache-matrix /
// cache-vector / baseline variants all run the one persistent kernel.
#pragma once

#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../perks_amd.h"

namespace perks_cg {

// The part of the command-line helper the reference's driver uses:
// "--name" flags and "--name=value" arguments.
class CommandLineArgs {
public:
    CommandLineArgs(int argc, char** argv)
    {
        for (int i = 1; i < argc; ++i) {
            std::string a(argv[i]);
            if (a.rfind("--", 0) != 0) continue;
            a = a.substr(2);
            const size_t eq = a.find('=');
            keys_.push_back(eq == std::string::npos ? a : a.substr(0, eq));
            values_.push_back(eq == std::string::npos ? std::string() : a.substr(eq + 1));
        }
    }
    bool CheckCmdLineFlag(const char* name) const
    {
        for (const std::string& k : keys_)
            if (k == name) return true;
        return false;
    }
    template <typename T>
    void GetCmdLineArgument(const char* name, T& val) const
    {
        for (size_t i = 0; i < keys_.size(); ++i)
            if (keys_[i] == name) {
                std::istringstream s(values_[i]);
                s >> val;
            }
    }

private:
    std::vector<std::string> keys_, values_;
};

}  // namespace perks_cg

using perks_cg::CommandLineArgs;

template <typename ValueT, typename OffsetT, bool baseline, bool cacheMatrix, bool cacheVector>
void myTest(int devID, perks_device_caps& deviceProp, CommandLineArgs& args)
{
    static_assert(sizeof(OffsetT) == 4, "32-bit offsets");
    static_assert(sizeof(ValueT) == 4 || sizeof(ValueT) == 8, "float or double values");
    (void)devID;
    (void)deviceProp;
    std::string mtx;
    args.GetCmdLineArgument("mtx", mtx);
    int max_iter = -1;
    args.GetCmdLineArgument("iters", max_iter);
    perks_cg_driver_opts o;
    std::memset(&o, 0, sizeof(o));
    o.mtx_path = args.CheckCmdLineFlag("vdata") ? "" : mtx.c_str();
    o.value_bytes = (int)sizeof(ValueT);
    o.max_iter = max_iter;
    o.static_iter = args.CheckCmdLineFlag("staticiter");
    o.warmup = args.CheckCmdLineFlag("warmup");
    o.check = args.CheckCmdLineFlag("check");
    o.verbose = 1;
    perks_cg_driver_result r;
    const int rc = perks_cg_driver(&o, &r);
    if (rc != PERKS_OK) std::fprintf(stderr, "perks_cg_driver: %s\n", perks_error_string(rc));
}
