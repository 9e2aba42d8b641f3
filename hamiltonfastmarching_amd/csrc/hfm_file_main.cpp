// hfm_file_main.cpp -- FileHFM-compatible executable over the C ABI, a drop-in for the
// reference's per-model executables (Interfaces/FileHFM/FileHFM.cpp:12-26) as called by
// JMM_CPPLibs/FileIO/FileIO.py WriteCallRead:
//
//     FileHFM_<Model> [inputPrefix = input] [outputPrefix = output]
//
// reads <inputPrefix>_Format.txt / _Data.dat, solves on the GPU, writes the computed keys to
// <outputPrefix>_Format.txt / _Data.dat.  Named FileHFM_<Model>, the executable applies to that
// model only (a 'model' key must then name it, Headers/DispatchAndRun.h:227-233); under any
// other name the 'model' key selects the model.  As in the reference, the output pair is written
// even when the run fails after the input was read, and failures print "Exception caught. ...".
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "hfm_amd.h"

int main(int argc, const char *argv[]) {
    const std::string inputPrefix = argc > 1 ? argv[1] : "input";
    const std::string outputPrefix = argc > 2 ? argv[2] : "output";
    std::string name = argv[0];
    const size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) name = name.substr(slash + 1);
    const std::string tag = "FileHFM_";
    std::string fixed;
    if (name.compare(0, tag.size(), tag) == 0) {
        fixed = name.substr(tag.size());
        if (fixed != "Riemann2" && fixed != "Riemann3" && fixed != "ReedsSheppForward2" && fixed != "Elastica2" &&
            fixed != "Elastica2_5")
            fixed.clear();
    }
    hfm_io *io = hfm_io_create();
    if (hfm_io_read_files(io, inputPrefix.c_str())) {
        std::fprintf(stderr, "Exception caught. %s\n", hfm_io_last_error(io));
        hfm_io_destroy(io);
        return EXIT_FAILURE;
    }
    int rc = 0;
    std::string msg;
    if (!fixed.empty()) {
        if (hfm_io_has_field(io, "model")) {
            std::vector<char> buf(256);
            const long long n = hfm_io_get_string(io, "model", buf.data(), buf.size());
            const std::string model = n >= 0 ? std::string(buf.data()) : std::string();
            if (model != fixed && !(fixed == "Elastica2" && model == "Elastica2_5")) {
                rc = 1;
                msg = "Executable applies to " + fixed + ", not to " + model + ".";
            }
        } else {
            hfm_io_set_string(io, "model", fixed.c_str());
        }
    }
    if (!rc) {
        rc = hfm_run(io);
        if (rc) msg = hfm_io_last_error(io);
    }
    const char *messages = hfm_io_messages(io);
    if (messages && *messages) std::fputs(messages, stdout);
    if (rc) std::fprintf(stderr, "Exception caught. %s\n", msg.c_str());
    if (hfm_io_write_files(io, outputPrefix.c_str())) {
        std::fprintf(stderr, "Exception caught. %s\n", hfm_io_last_error(io));
        rc = 1;
    }
    hfm_io_destroy(io);
    return rc ? EXIT_FAILURE : EXIT_SUCCESS;
}
