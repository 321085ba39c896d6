// main.cpp -- command line with the reference's shape (src/main.cpp:51-70):
//   edge_matching_puzzle <definition.txt> <ressources_dir> <feature_name>
//                        [--strategy=basic|spiral] [--hint=<depth>]
//                        [--enable_gui=false] [--dump_file=] [--initial_situation=]
// new_strategy uses the spiral order and CUDA_backtracker (and
// CUDA_enumerate, situation_profile) the basic order unless --strategy is
// given
// (src/main.cpp:108-119); the dump file defaults to results.bin (:74).
// There is no GUI: --enable_gui=true is refused.  Errors print
// "ERROR : <what> at <file>:<line>" and return -1 (src/main.cpp:137-146).
#include <iostream>
#include <string>

#include "emp_hip_session.h"
#include "factory_feature.h"

using namespace edge_matching_puzzle;

int main(int argc, char **argv) {
    try {
        if (argc < 4) {
            std::cerr << "Usage: " << argv[0]
                      << " <definition.txt> <ressources_dir> <feature_name> [--strategy=] [--hint=]" << std::endl;
            return -1;
        }
        std::string strategy, hint, initial_situation, dump_file;
        bool dump_file_set = false;
        for (int i = 4; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--strategy=", 0) == 0) strategy = a.substr(11);
            else if (a.rfind("--hint=", 0) == 0) hint = a.substr(7);
            else if (a.rfind("--initial_situation=", 0) == 0) initial_situation = a.substr(20);
            else if (a.rfind("--dump_file=", 0) == 0) dump_file = a.substr(12), dump_file_set = true;
            else if (a.rfind("--enable_gui=", 0) == 0) {
                const std::string v = a.substr(13);
                if (v != "false" && v != "0") throw logic_exception("this build has no GUI: use --enable_gui=false", __LINE__, __FILE__);
            } else throw logic_exception("Unknown option \"" + a + "\"", __LINE__, __FILE__);
        }
        const std::string feature = argv[3];
        if (strategy.empty())
            strategy = feature == "CUDA_backtracker" || feature == "CUDA_enumerate" || feature == "situation_profile"
                           ? "basic"
                           : "spiral";
        if (dump_file_set && feature != "new_strategy")
            throw logic_exception("--dump_file is only written by new_strategy", __LINE__, __FILE__);
        if (!dump_file_set) dump_file = "results.bin";
        puzzle_description puzzle = parse_puzzle_file(argv[1]);
        std::cout << puzzle.width * puzzle.height << " pieces loaded" << std::endl;
        emp_hip_puzzle view = puzzle.view();
        EMP_HIP_CALL(emp_hip_validate_puzzle(&view));
        std::unique_ptr<feature_if> f =
            factory_feature::create_feature(feature, puzzle, initial_situation, dump_file, strategy, hint);
        f->run();
    } catch (const logic_exception &e) {
        std::cout << "ERROR : " << e.what() << " at " << e.file() << ":" << e.line() << std::endl;
        return -1;
    } catch (const std::exception &e) {
        std::cout << "ERROR : " << e.what() << std::endl;
        return -1;
    }
    return 0;
}
