# find_package(Tensorforge) / find_package(tensorforge): CMake looks for
# <lower-case-name>-config.cmake under any spelling of the package name (the reference's own
# integration test spells it Tensorforge). Same package as TensorForgeConfig.cmake.
include("${CMAKE_CURRENT_LIST_DIR}/TensorForgeConfig.cmake")
