#pragma once
#include "SourceStage.h"
