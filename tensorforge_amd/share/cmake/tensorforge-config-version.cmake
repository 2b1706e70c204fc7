include("${CMAKE_CURRENT_LIST_DIR}/TensorForgeConfigVersion.cmake")
