/*
 * mcx_presets.c - built-in benchmark inputs selected with --bench / -Q.
 *
 * The inputs are the reference's benchmark definitions (src/mcx_bench.h:31-560):
 * same domains, media, sources, detectors, seeds and photon counts, so that
 * published absorbed fractions (speedcontest/mcxcontest:103-105,
 * test/testmcx.sh:61-113) can be checked against them.
 */
#include <string.h>

#define MCX_JSON(...) #__VA_ARGS__

static const char* preset_names[] = {"cube60", "cube60b", "cube60planar", "cubesph60b", "onecube", "twocube",
                                     "skinvessel", "sphshells", NULL
                                    };

const char* benchname[] = {"cube60", "cube60b", "cube60planar", "cubesph60b", "onecube", "twocube", "skinvessel",
                           "sphshells", ""
                          };

#define CUBE60_MEDIA \
    "\"Media\":[{\"mua\":0,\"mus\":0,\"g\":1,\"n\":1},{\"mua\":0.005,\"mus\":1.0,\"g\":0.01,\"n\":1.37}," \
    "{\"mua\":0.002,\"mus\":5.0,\"g\":0.90,\"n\":1.0}]"
#define CUBE60_DETS \
    "\"Detector\":[{\"Pos\":[29,19,0],\"R\":1},{\"Pos\":[29,39,0],\"R\":1},{\"Pos\":[19,29,0],\"R\":1}," \
    "{\"Pos\":[39,29,0],\"R\":1}]"
#define CUBE60_HEAD(id, mismatch) \
    "{\"Session\":{\"ID\":\"" id "\",\"Photons\":1e6,\"RNGSeed\":1648335518,\"DoMismatch\":" mismatch "}," \
    "\"Domain\":{\"Dim\":[60,60,60],\"OriginType\":1," CUBE60_MEDIA "}," \
    "\"Forward\":{\"T0\":0,\"T1\":5e-9,\"Dt\":5e-9},"

static const char* preset_json[] = {
    CUBE60_HEAD("cube60", "0")
    "\"Shapes\":[{\"Name\":\"cubic60\"},{\"Origin\":[0,0,0]},{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}}],"
    "\"Optode\":{\"Source\":{\"Type\":\"pencil\",\"Pos\":[29,29,0],\"Dir\":[0,0,1]}," CUBE60_DETS "}}",

    CUBE60_HEAD("cube60b", "true")
    "\"Shapes\":[{\"Name\":\"cube60b\"},{\"Origin\":[0,0,0]},{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}}],"
    "\"Optode\":{\"Source\":{\"Type\":\"pencil\",\"Pos\":[29,29,0],\"Dir\":[0,0,1]}," CUBE60_DETS "}}",

    CUBE60_HEAD("cube60planar", "true")
    "\"Shapes\":[{\"Name\":\"cube60planar\"},{\"Origin\":[0,0,0]},{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}}],"
    "\"Optode\":{\"Source\":{\"Type\":\"planar\",\"Pos\":[10,10,-10],\"Dir\":[0,0,1],"
    "\"Param1\":[40,0,0,0],\"Param2\":[0,40,0,0]}," CUBE60_DETS "}}",

    CUBE60_HEAD("cubesph60b", "true")
    "\"Shapes\":[{\"Name\":\"cubesph60b\"},{\"Origin\":[0,0,0]},{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}},"
    "{\"Sphere\":{\"Tag\":2,\"O\":[30,30,30],\"R\":15}}],"
    "\"Optode\":{\"Source\":{\"Type\":\"pencil\",\"Pos\":[29,29,0],\"Dir\":[0,0,1]}," CUBE60_DETS "}}",

    MCX_JSON({"Session": {"ID": "onecube", "DoSaveVolume": 0, "DoAutoThread": 1, "SaveDataMask": "xv", "Photons": 1000000},
              "Forward": {"T0": 0, "T1": 5e-09, "Dt": 5e-09},
              "Optode": {"Source": {"Pos": [0.5, 0.5, 0], "Dir": [0, 0, 1]},
                         "Detector": [{"Pos": [0.5, 0.45, 0], "R": 0.0166666667}]},
              "Domain": {"OriginType": 1, "LengthUnit": 60,
                         "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.005, "mus": 1, "g": 0, "n": 1.37}],
                         "MediaFormat": "byte", "Dim": [1, 1, 1]},
              "Shapes": {"_ArraySize_": [1, 1, 1], "_ArrayType_": "uint8", "_ArrayData_": [1]}}),

    MCX_JSON({"Session": {"ID": "twocube", "DoSaveVolume": 0, "DoAutoThread": 1, "SaveDataMask": "xv", "Photons": 1000000},
              "Forward": {"T0": 0, "T1": 5e-09, "Dt": 5e-09},
              "Optode": {"Source": {"Pos": [1, 1, 0], "Dir": [0, 0, 1]},
                         "Detector": [{"Pos": [1, 0.9, 0], "R": 0.033333333333}]},
              "Domain": {"OriginType": 1, "LengthUnit": 30,
                         "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.005, "mus": 1, "g": 0, "n": 1.37}],
                         "MediaFormat": "byte", "Dim": [2, 2, 2]},
              "Shapes": {"_ArraySize_": [2, 2, 2], "_ArrayType_": "uint8", "_ArrayData_": [1, 1, 1, 1, 1, 1, 1, 1]}}),

    MCX_JSON({"Session": {"ID": "skinvessel", "Photons": 1e6, "RNGSeed": 1648335518, "DoMismatch": true, "DoAutoThread": true},
              "Forward": {"T0": 0, "T1": 5e-08, "Dt": 5e-08},
              "Optode": {"Source": {"Pos": [100, 100, 20], "Dir": [0, 0, 1], "Param1": [60, 0, 0, 0], "Type": "disk"}},
              "Domain": {"OriginType": 1, "LengthUnit": 0.005,
                         "Media": [{"mua": 1e-05, "mus": 0, "g": 1, "n": 1.37},
                                   {"mua": 3.564e-05, "mus": 1, "g": 1, "n": 1.37},
                                   {"mua": 23.05426549, "mus": 9.398496241, "g": 0.9, "n": 1.37},
                                   {"mua": 0.04584957865, "mus": 35.65405549, "g": 0.9, "n": 1.37},
                                   {"mua": 1.657237447, "mus": 37.59398496, "g": 0.9, "n": 1.37}],
                         "Dim": [200, 200, 200]},
              "Shapes": [{"Grid": {"Tag": 1, "Size": [200, 200, 200]}},
                         {"ZLayers": [[1, 20, 1], [21, 32, 4], [33, 200, 3]]},
                         {"Cylinder": {"Tag": 2, "C0": [0, 100.5, 100.5], "C1": [200, 100.5, 100.5], "R": 20}}]}),

    MCX_JSON({"Session": {"ID": "sphshells", "Photons": 1e6, "RNGSeed": 1648335518, "DoMismatch": true,
                          "DoPartialPath": false, "DoAutoThread": true},
              "Forward": {"T0": 0, "T1": 5e-09, "Dt": 5e-09},
              "Domain": {"LengthUnit": 1,
                         "Media": [{"mua": 0.0, "mus": 0.0, "g": 1.0, "n": 1.0}, {"mua": 0.02, "mus": 7.0, "g": 0.89, "n": 1.37},
                                   {"mua": 0.004, "mus": 0.009, "g": 0.89, "n": 1.37}, {"mua": 0.02, "mus": 9.0, "g": 0.89, "n": 1.37},
                                   {"mua": 0.05, "mus": 0.0, "g": 1.0, "n": 1.37}],
                         "Dim": [60, 60, 60], "OriginType": 1},
              "Optode": {"Source": {"Type": "pencil", "Pos": [30, 30.1, 0], "Dir": [0, 0, 1, 0], "SrcNum": 1}},
              "Shapes": [{"Grid": {"Tag": 1, "Size": [60, 60, 60]}}, {"Sphere": {"Tag": 2, "O": [30, 30, 30], "R": 25}},
                         {"Sphere": {"Tag": 3, "O": [30, 30, 30], "R": 23}}, {"Sphere": {"Tag": 4, "O": [30, 30, 30], "R": 10}}]}),
};

const char* mcx_preset_json(const char* name) {
    for (int i = 0; preset_names[i]; i++) {
        if (!strcmp(name, preset_names[i])) {
            return preset_json[i];
        }
    }

    return NULL;
}
