#pragma once
#include <dt/neighborhood.hpp>
