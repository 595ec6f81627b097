// atrc_cli: Atrc's command line (src/cli/src/main.cpp:14-117, params.cpp:6-43)
//   atrc_cli -d scene.json        render every session of the scene file
//   atrc_cli -s "<json>"          scene description given inline
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "factory.h"
#include "renderer.h"

int main(int argc, char **argv) {
    std::string text, filename = "./scene.txt";
    bool have_text = false, have_file = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if ((a == "-s" || a == "--scene") && i + 1 < argc) { text = argv[++i]; have_text = true; }
        else if ((a == "-d" || a == "--scene-filename") && i + 1 < argc) { filename = argv[++i]; have_file = true; }
        else if (a == "-h" || a == "--help") {
            std::cout << "atrc_cli -d scene.json | -s \"<json>\"\n";
            return 0;
        }
    }
    try {
        if (!have_text) {
            if (!have_file) throw std::runtime_error("scene description is unspecified");
            std::ifstream in(filename);
            if (!in) throw std::runtime_error("cannot read " + filename);
            std::stringstream ss;
            ss << in.rdbuf();
            text = ss.str();
        }
        atrc::PathMapper mapper;
        mapper.add_replacer("${working-directory}",
                            std::filesystem::absolute(std::filesystem::current_path()).lexically_normal().string());
        mapper.add_replacer("${scene-directory}",
                            std::filesystem::absolute(std::filesystem::path(filename)).parent_path().lexically_normal().string());
        atrc::CreatingContext ctx;
        ctx.path_mapper = &mapper;
        auto root = atrc::json_to_config(text);
        auto sessions = atrc::parse_render_sessions(*root, ctx);
        for (size_t i = 0; i < sessions.size(); ++i) {
            std::cout << "processing rendering session [" << i << "]" << std::endl;
            sessions[i].execute();
        }
        return 0;
    } catch (const std::exception &e) {
        std::cout << e.what() << std::endl;
    }
    return -1;
}
