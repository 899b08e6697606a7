// The create/destroy pair of one drop-in plugin object (liblpmd/lpmd/plugin.h:42-46).
// Compiled once per reference plugin name with -DLPMDX_PLUGIN_NAME='"<name>"'
// into <name>.so, each linked against liblpmdx_lpmd.so.
#include "lpmdx_lpmd.hpp"

#ifndef LPMDX_PLUGIN_NAME
#error "build with -DLPMDX_PLUGIN_NAME='\"<reference plugin name>\"'"
#endif

// No exception leaves create(): an error ends the program with its message
// (lpmdx_lpmd::CreateOrExit).
lpmd::Plugin* create(std::string args) { return lpmdx_lpmd::CreateOrExit(LPMDX_PLUGIN_NAME, args); }
void destroy(lpmd::Plugin* m) { delete m; }
