// Definition of the opaque host scene handle shared by the host and device
// halves of the C ABI.
#pragma once
#include <string>

#include "host/flat.h"
#include "host/scene.h"

struct xt_scene {
    xt::Scene scene;
    xt::FlatScene flat;
    bool built = false;
    std::string error;
};
