// Forwarding header: the reference's <AiLBFlow.h> resolved to the reference-API layer
// (mkshim.h over include/mkg.h), so that MaterKit-style drivers compile unchanged with
// -I<this folder> -I<shim> -I<include> and link against libmkshim.so and libmkg.so.  Like
// the reference's headers, it brings the std and mkt namespaces into scope.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "mkshim.h"

using namespace std;
using namespace mkt;
