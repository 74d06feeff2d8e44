#pragma once
#include "Port.h"
