// parser.h -- "--key=value" command-line parser of the benchmark programs
// (reference src/util/parser.h:7-97; keys and defaults are the same).
#pragma once
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_map>

namespace MGBrain {

class ParamsParser {
   private:
    std::unordered_map<std::string, std::string> params;

   public:
    inline static std::string NSYN = "--nsyn";
    inline static std::string NPART = "--npart";
    inline static std::string MODEL = "--model";
    inline static std::string FILE = "--file";
    inline static std::string PREFIX = "--prefix";
    inline static std::string RATE = "--rate";

    ParamsParser(int args, char* argv[]) {
        params[NSYN] = "10000";
        params[NPART] = "1";
        params[MODEL] = "model";
        params[FILE] = "tmp.txt";
        params[PREFIX] = "../../benchdata/";
        params[RATE] = "0.002";
        for (int i = 1; i < args; i++) {
            const std::string arg(argv[i]);
            const size_t eq = arg.find('=');
            if (eq == std::string::npos) params[arg] = "";
            else params[arg.substr(0, eq)] = arg.substr(eq + 1);
        }
    }
    std::string getStr(std::string name) { return params.count(name) ? params[name] : ""; }
    size_t getNum(std::string name) {
        if (!params.count(name)) return (size_t)-1;
        std::stringstream ss(params[name]);
        size_t res = 0;
        ss >> res;
        return res;
    }
    float getFloat(std::string name) {
        if (!params.count(name)) return -1;
        std::stringstream ss(params[name]);
        float res = 0;
        ss >> res;
        return res;
    }
    bool getBool(std::string name) { return params.count(name) && params[name] == "y"; }
};

}  // namespace MGBrain
