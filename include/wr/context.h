#pragma once
// libWR/include/wr/context.h:1-31 — same names, fields and defaults.
#include <cstdint>

namespace wr {

class Scene;
class Renderer;

struct ContextConfig {
    int deviceId = 0;
    int logLevel = 4;  // 0 silences the [WR]/[Scene]/[Renderer] log lines
    bool enableValidation = false;
};

// Process-wide device context (HIP device selection and initialisation).
class Context {
public:
    Context(const ContextConfig& config = {});
    ~Context();

    Scene* createScene();
    Renderer* createRenderer();

    int getDeviceId() const;
};

}  // namespace wr
