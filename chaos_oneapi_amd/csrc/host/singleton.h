// Meyers singleton, as the reference's singleton.h:4-12.
#pragma once

template <typename T>
class Singleton {
public:
    static T& Instance() {
        static T _instance;
        return _instance;
    }
};
